// tp_hip.hip -- libtp: gfx950 kernels and the C ABI (include/tp.h) for the
// ThreePhotons.jl three-photon correlation c3 and its Bayes energy.
//
// Evaluation of one energy (lambda = 0) is three launches on the context
// stream:
//   k_triple      coupling-form triple products As[lt][kt] (FP64 VALU/LDS)
//   k_contract    c3 = F . As on FP64 MFMA (v_mfma_f64_16x16x4f64), fused
//                 Bayes epilogue: per (row tile, column) partial sums of
//                 p = max(c3,1e-30) and qhat*log(p)
//   k_final       per-column energies and a fixed-order sum (deterministic)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/tp.h"
#include "tp_host_math.hpp"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int tp_fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define TP_HIP(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return tp_fail(TP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
        } while (0)

// ---------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------
struct tp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double last_ms[3] = {0, 0, 0};
};

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

template <class T>
static int dev_alloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) return tp_fail(TP_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

// Tile shape of the FP64 contraction.
constexpr int CT_BM = 128;  // rows (alpha,beta bins) per workgroup
constexpr int CT_BN = 64;   // columns (k-triples) per workgroup
constexpr int CT_BK = 16;   // l-triples per LDS stage

struct tp_basis {
    tp_ctx *ctx = nullptr;
    int L = 0, LMAX = 0, N = 0, K = 0, for_intensity = 1, lstep = 2;
    double lambda = 0, dq = 0;
    int64_t basislen = 0, n_lt = 0, klength = 0, R = 0;
    int n_c = 0;                 // compact coefficients per shell (l in lrange, m = -l..l)
    std::vector<int> coff;       // compact offset of order l (indexed by l)
    std::vector<int> sean0;      // [n_c] seanindex-1 of each compact entry
    int64_t R_pad = 0, nlt_pad = 0, kl_pad = 0;
    int n_rt = 0, n_ct = 0;      // row / column tiles of the contraction
    int n_groups = 0, nchunks = 0, max_chunk_groups = 0, npairs = 0;
    // device tables
    double *d_F = nullptr;         // [nlt_pad][R_pad]
    double *d_ew = nullptr;        // [basislen] weights, sorted (lt, mu3, m1)
    uint32_t *d_eidx = nullptr;    // [basislen] compact i1 | i2 << 16
    int *d_gstart = nullptr;       // [n_groups + 1]
    int4 *d_ltinfo = nullptr;      // [n_lt] {coff[l3], goff, l3, 0}
    int4 *d_chunks = nullptr;      // [nchunks] {lt0, lt1, g0, g1}
    int *d_pairs = nullptr;        // [npairs] k1 | k2 << 16 (1-based)
    int *d_sean0 = nullptr;        // [n_c]
    // per-evaluation scratch
    double *d_As = nullptr;        // [nlt_pad][kl_pad]
    double *d_part = nullptr;      // [n_rt][2][kl_pad]
    double *d_out = nullptr;       // [8] scalar results
    double2 *d_raw = nullptr;      // [K][LMAX^2] staging of host coefficients
    double2 *d_coef = nullptr;     // [K][n_c] compact coefficients of host-input calls
    double *d_c3 = nullptr;        // [klength][R] c3raw output staging (lazy)
};

struct tp_ref {
    const tp_basis *b = nullptr;
    int64_t ncols = 0;
    double *d_qhat = nullptr;  // [ncols][R]
    double *d_qsum = nullptr;  // [ncols]
    double *d_tf = nullptr;    // [ncols]
};

struct tp_state {
    const tp_basis *b = nullptr;
    int nshell = 0;
    double2 *d_cur = nullptr;    // [nshell][n_c]
    double2 *d_prop = nullptr;   // [nshell][n_c]
    double *d_scal = nullptr;    // [8]: E, T, ...
    double *d_step = nullptr;    // [L/2] step sizes
    int *d_flag = nullptr;       // [1] accept flag
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller normals.  Streams are
// keyed by the run seed; the counter is (j>>1, l-index, step, tag).
// ---------------------------------------------------------------------------
#define TP_TAG_MOVE 0x4D4F5645u
#define TP_TAG_ACC 0x41434350u
#define TP_TAG_TEMP 0x54454D50u

__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

__device__ __forceinline__ double u01(uint32_t lo, uint32_t hi)
{
    const uint64_t u = ((uint64_t)hi << 32) | lo;
    return (double)(u >> 11) * (1.0 / 9007199254740992.0);
}

__device__ double dev_normal(uint64_t seed, uint32_t step, uint32_t li, uint32_t j, uint32_t tag)
{
    uint32_t c[4] = {j >> 1, li, step, tag};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double a = u01(c[0], c[1]), b = u01(c[2], c[3]);
    const double r = sqrt(-2.0 * log(1.0 - a));
    const double ang = 2.0 * 3.14159265358979323846 * b;
    return (j & 1) ? r * sin(ang) : r * cos(ang);
}

__device__ double dev_uniform(uint64_t seed, uint32_t step, uint32_t tag)
{
    uint32_t c[4] = {0u, 0xFFFFFFFFu, step, tag};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return u01(c[0], c[1]);
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Gather the compact (even l <= L) coefficients out of LMAX^2-strided vectors.
__global__ void k_pack(const double2 *__restrict__ raw, int64_t ld, int nshell, const int *__restrict__ sean0,
                       int n_c, double2 *__restrict__ coef)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nshell * n_c) return;
    const int k = (int)(i / n_c), c = (int)(i % n_c);
    coef[i] = raw[(int64_t)k * ld + sean0[c]];
}

// Triple products in coupling form.  Workgroup (pair p = (k1,k2), chunk ch):
//   stage 1: T[lt,mu3] = sum_{m1} w(l1 m1 l2 m2 l3 mu3) c_k1[l1,m1] c_k2[l2,m2],
//            m2 = -m1-mu3, for the (lt,mu3) groups of the chunk (LDS);
//   stage 2: As[lt][kt(k1,k2,k3)] = Re sum_mu3 T[lt,mu3] c_k3[l3,mu3], k3 <= k2.
// This equals sum_{j in lt} w_j Re(c_k1[i1] c_k2[i2] c_k3[i3])
// (correlations.jl:165-174, cuda_kernel.cu:30-37) regrouped by mu3.
__global__ __launch_bounds__(256) void k_triple(const double2 *__restrict__ coef, int n_c,
                                                const double *__restrict__ ew, const uint32_t *__restrict__ eidx,
                                                const int *__restrict__ gstart, const int4 *__restrict__ ltinfo,
                                                const int4 *__restrict__ chunks, int nchunks,
                                                const int *__restrict__ pairs, double *__restrict__ As,
                                                int64_t kl_pad)
{
    extern __shared__ double2 sm[];
    double2 *c1 = sm, *c2 = sm + n_c, *T = sm + 2 * n_c;
    const int p = blockIdx.x / nchunks, ch = blockIdx.x % nchunks;
    const int pk = pairs[p], k1 = pk & 0xffff, k2 = pk >> 16;
    for (int i = threadIdx.x; i < n_c; i += blockDim.x) {
        c1[i] = coef[(int64_t)(k1 - 1) * n_c + i];
        c2[i] = coef[(int64_t)(k2 - 1) * n_c + i];
    }
    __syncthreads();
    const int4 cw = chunks[ch];
    const int lt0 = cw.x, lt1 = cw.y, g0 = cw.z, g1 = cw.w;
    for (int g = g0 + threadIdx.x; g < g1; g += blockDim.x) {
        const int s = gstart[g], e = gstart[g + 1];
        double tr = 0.0, ti = 0.0;
        for (int j = s; j < e; ++j) {
            const double w = ew[j];
            const uint32_t id = eidx[j];
            const double2 a = c1[id & 0xffffu], b = c2[id >> 16];
            tr += w * (a.x * b.x - a.y * b.y);
            ti += w * (a.x * b.y + a.y * b.x);
        }
        T[g - g0] = make_double2(tr, ti);
    }
    __syncthreads();
    const int nlt = lt1 - lt0;
    const int64_t base = (int64_t)(k1 - 1) * k1 * (k1 + 1) / 6 + (int64_t)(k2 - 1) * k2 / 2;
    for (int o = threadIdx.x; o < k2 * nlt; o += blockDim.x) {
        const int k3 = o % k2 + 1, lt = lt0 + o / k2;
        const int4 li = ltinfo[lt];
        const double2 *c3 = coef + (int64_t)(k3 - 1) * n_c + li.x;
        const double2 *Tl = T + (li.y - g0);
        const int n = 2 * li.z + 1;
        double s = 0.0;
        for (int m = 0; m < n; ++m) {
            const double2 t = Tl[m], c = c3[m];
            s += t.x * c.x - t.y * c.y;
        }
        As[(int64_t)lt * kl_pad + base + k3 - 1] = s;
    }
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// c3[r,kt] = sum_lt F[r,lt] As[lt,kt] on FP64 MFMA, BM x BN tile per
// workgroup of 4 waves (each wave BM/4 rows x BN columns), l-triples staged
// through LDS in slices of 16.  Epilogue: raw c3 out (optional) and/or the
// Bayes partial sums over the tile's rows (energy, correlations.jl:203-220).
// Workgroups are mapped so that each XCD (block id mod 8) owns whole row
// tiles: its F rows stay in that XCD's L2 while As streams through.
template <bool WRITE_C3, bool ENERGY>
__global__ __launch_bounds__(256) void k_contract(const double *__restrict__ F, int64_t R_pad,
                                                  const double *__restrict__ As, int64_t kl_pad, int nlt_pad,
                                                  int R, int ncols, int n_rt, int n_ct,
                                                  const double *__restrict__ qhat, double *__restrict__ c3out,
                                                  double *__restrict__ part)
{
    constexpr int BM = CT_BM, BN = CT_BN, BK = CT_BK;
    constexpr int WM = BM / 4, RB = WM / 16, CB = BN / 16;
    __shared__ double sA[BK][BM];
    __shared__ double sB[BK][BN];
    __shared__ double red[4][2][BN];

    const int nblk = n_rt * n_ct;
    int lin = blockIdx.x;
    if ((nblk & 7) == 0) lin = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
    const int rt = lin / n_ct, ct = lin % n_ct;
    const int r0 = rt * BM, c0 = ct * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;

    dbl4 acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    // global -> register prefetch of one BK slice
    constexpr int A2 = BK * BM / 2 / 256;   // double2 per thread for A
    constexpr int B2 = BK * BN / 2 / 256;   // double2 per thread for B
    double2 ra[A2], rb[B2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < A2; ++q) {
            const int i = tid + 256 * q, row = i / (BM / 2), c2 = i % (BM / 2);
            ra[q] = *reinterpret_cast<const double2 *>(F + (int64_t)(k0 + row) * R_pad + r0 + 2 * c2);
        }
#pragma unroll
        for (int q = 0; q < B2; ++q) {
            const int i = tid + 256 * q, row = i / (BN / 2), c2 = i % (BN / 2);
            rb[q] = *reinterpret_cast<const double2 *>(As + (int64_t)(k0 + row) * kl_pad + c0 + 2 * c2);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int q = 0; q < A2; ++q) {
            const int i = tid + 256 * q, row = i / (BM / 2), c2 = i % (BM / 2);
            *reinterpret_cast<double2 *>(&sA[row][2 * c2]) = ra[q];
        }
#pragma unroll
        for (int q = 0; q < B2; ++q) {
            const int i = tid + 256 * q, row = i / (BN / 2), c2 = i % (BN / 2);
            *reinterpret_cast<double2 *>(&sB[row][2 * c2]) = rb[q];
        }
    };

    gload(0);
    for (int k0 = 0; k0 < nlt_pad; k0 += BK) {
        __syncthreads();
        sstore();
        __syncthreads();
        if (k0 + BK < nlt_pad) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            double a[RB], b[CB];
#pragma unroll
            for (int i = 0; i < RB; ++i) a[i] = sA[kk * 4 + lq][wave * WM + i * 16 + l16];
#pragma unroll
            for (int j = 0; j < CB; ++j) b[j] = sB[kk * 4 + lq][j * 16 + l16];
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }

    // accumulator layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane
    // holds column l16, rows lq + 4*v (v = 0..3) of the 16x16 block.
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const int col = c0 + j * 16 + l16;
        const bool cok = col < ncols;
        double sp = 0.0, sq = 0.0;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int rbase = r0 + wave * WM + i * 16 + lq;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rbase + 4 * v;
                const double c = acc[i][j][v];
                if (cok && r < R) {
                    if (WRITE_C3) c3out[(int64_t)col * R + r] = c;
                    if (ENERGY) {
                        const double pv = c > 1e-30 ? c : 1e-30;
                        const double q = qhat[(int64_t)col * R + r];
                        sp += pv;
                        sq += q * log(pv);
                    }
                }
            }
        }
        if (ENERGY) {
            sp += __shfl_xor(sp, 16);
            sq += __shfl_xor(sq, 16);
            sp += __shfl_xor(sp, 32);
            sq += __shfl_xor(sq, 32);
            if (lq == 0) {
                red[wave][0][j * 16 + l16] = sp;
                red[wave][1][j * 16 + l16] = sq;
            }
        }
    }
    if (ENERGY) {
        __syncthreads();
        if (tid < BN) {
            const double sp = ((red[0][0][tid] + red[1][0][tid]) + red[2][0][tid]) + red[3][0][tid];
            const double sq = ((red[0][1][tid] + red[1][1][tid]) + red[2][1][tid]) + red[3][1][tid];
            part[((int64_t)rt * 2 + 0) * kl_pad + c0 + tid] = sp;
            part[((int64_t)rt * 2 + 1) * kl_pad + c0 + tid] = sq;
        }
    }
}

// E = sum_kt tF * (qsum * log(S) - Q), S = sum_r p, Q = sum_r qhat log p.
// One workgroup of 1024 threads; fixed assignment and a fixed tree, so the
// result is bitwise reproducible.
__global__ __launch_bounds__(1024) void k_final(const double *__restrict__ part, int n_rt, int64_t kl_pad,
                                                int ncols, const double *__restrict__ qsum,
                                                const double *__restrict__ tf, double *__restrict__ out)
{
    __shared__ double red[1024];
    double e = 0.0;
    for (int kt = threadIdx.x; kt < ncols; kt += blockDim.x) {
        double S = 0.0, Q = 0.0;
        for (int rt = 0; rt < n_rt; ++rt) {
            S += part[((int64_t)rt * 2 + 0) * kl_pad + kt];
            Q += part[((int64_t)rt * 2 + 1) * kl_pad + kt];
        }
        e += tf[kt] * (qsum[kt] * log(S) - Q);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// One Monte Carlo proposal (get_MC_step, determination.jl:232-241): for each
// even l in 2..L (one workgroup per l) a rotation by beta = stepsize_l * z in
// the plane of a random orthonormal 2-frame (u,v) of R^(2l+1),
//   R = I + (cos b - 1)(u u^T + v v^T) + sin b (v u^T - u v^T),
// which is inv(C) G(beta) C of random_rotation_step (utilities.jl:180-190)
// written through the two rows of C that G touches.  It acts on the complex
// vectors through the real basis, x <- ctr^H R ctr x, with ctr = Umat(l)
// (spherical_harmonics.jl:445-457) applied in its two-entries-per-row form.
// Workgroup 0 also copies the l = 0 block.
__global__ __launch_bounds__(64) void k_propose(const double2 *__restrict__ cur, double2 *__restrict__ prop,
                                                int nshell, int n_c, const int *__restrict__ coff_l,
                                                const double *__restrict__ stepsizes, uint64_t seed,
                                                uint32_t step, uint32_t tag)
{
    __shared__ double g[2 * 37 + 1];
    __shared__ double uv[2][37];
    __shared__ double sbeta;
    const int li = blockIdx.x, l = 2 * (li + 1), n = 2 * l + 1;
    const int tid = threadIdx.x;
    for (int j = tid; j < 2 * n + 1; j += blockDim.x) g[j] = dev_normal(seed, step, (uint32_t)li, (uint32_t)j, tag);
    __syncthreads();
    if (tid == 0) {
        double nu = 0.0;
        for (int i = 0; i < n; ++i) nu += g[i] * g[i];
        nu = sqrt(nu);
        for (int i = 0; i < n; ++i) uv[0][i] = g[i] / nu;
        double d = 0.0;
        for (int i = 0; i < n; ++i) d += uv[0][i] * g[n + i];
        double nv = 0.0;
        for (int i = 0; i < n; ++i) {
            const double t = g[n + i] - d * uv[0][i];
            uv[1][i] = t;
            nv += t * t;
        }
        nv = sqrt(nv);
        for (int i = 0; i < n; ++i) uv[1][i] /= nv;
        sbeta = stepsizes[li] * g[2 * n];
    }
    __syncthreads();
    const double cm = cos(sbeta) - 1.0, sb = sin(sbeta);
    const double h = 0.70710678118654752440;
    const int off = coff_l[l];
    for (int k = tid; k < nshell; k += blockDim.x) {
        const double2 *x = cur + (int64_t)k * n_c + off;
        double2 *xo = prop + (int64_t)k * n_c + off;
        if (li == 0) prop[(int64_t)k * n_c + coff_l[0]] = cur[(int64_t)k * n_c + coff_l[0]];
        // y = ctr x, index u + l
        double2 y[37];
        y[l] = x[l];
        for (int u = 1; u <= l; ++u) {
            const double sg = (u & 1) ? -1.0 : 1.0;
            const double2 xp = x[l + u], xm = x[l - u];
            y[l + u] = make_double2(h * (xp.x + sg * xm.x), h * (xp.y + sg * xm.y));
            // y_{-u} = (i (-1)^u x_{-u} - i x_u)/sqrt2
            y[l - u] = make_double2(h * (-sg * xm.y + xp.y), h * (sg * xm.x - xp.x));
        }
        double ar = 0, ai = 0, br = 0, bi = 0;
        for (int i = 0; i < n; ++i) {
            ar += uv[0][i] * y[i].x; ai += uv[0][i] * y[i].y;
            br += uv[1][i] * y[i].x; bi += uv[1][i] * y[i].y;
        }
        for (int i = 0; i < n; ++i) {
            const double u = uv[0][i], v = uv[1][i];
            y[i].x += cm * (u * ar + v * br) + sb * (v * ar - u * br);
            y[i].y += cm * (u * ai + v * bi) + sb * (v * ai - u * bi);
        }
        // x = ctr^H y
        xo[l] = y[l];
        for (int m = 1; m <= l; ++m) {
            const double sg = (m & 1) ? -1.0 : 1.0;
            const double2 yp = y[l + m], ym = y[l - m];
            // x_m = (y_m + i y_{-m})/sqrt2 ; x_{-m} = (-1)^m (y_m - i y_{-m})/sqrt2
            xo[l + m] = make_double2(h * (yp.x - ym.y), h * (yp.y + ym.x));
            xo[l - m] = make_double2(h * sg * (yp.x + ym.y), h * sg * (yp.y - ym.x));
        }
    }
}

// evaluate_step (determination.jl:491-515) with T *= decay first
// (rotate_all_at_once, :328): Metropolis decision on the exact energy,
// step-size update, traces.  scal = {E, T}.
__global__ void k_metropolis(double *__restrict__ scal, double *__restrict__ stepsizes, int nl,
                             const double *__restrict__ Enew, uint64_t seed, uint32_t step, double decay,
                             double factor, int *__restrict__ flag, double *__restrict__ trE,
                             double *__restrict__ trEn, int *__restrict__ trA)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double T = scal[1] * decay;
    scal[1] = T;
    const double En = Enew[0], E = scal[0];
    const double dE = En - E;
    const double ap = exp(-dE / T);
    const double u = dev_uniform(seed, step, TP_TAG_ACC);
    const int acc = (dE < 0.0 || (dE > 0.0 && u < ap)) ? 1 : 0;
    if (acc) {
        for (int i = 0; i < nl; ++i) stepsizes[i] *= factor;
        scal[0] = En;
    } else {
        for (int i = 0; i < nl; ++i) stepsizes[i] /= factor;
    }
    flag[0] = acc;
    trE[step - 1] = scal[0];
    trEn[step - 1] = En;
    trA[step - 1] = acc;
}

__global__ void k_accept_copy(const int *__restrict__ flag, const double2 *__restrict__ prop,
                              double2 *__restrict__ cur, int64_t n)
{
    if (flag[0] == 0) return;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        cur[i] = prop[i];
}

__global__ void k_store_scalar(const double *__restrict__ src, double *__restrict__ dst, int idx)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) dst[idx] = src[0];
}

// ---- synthetic intensity --------------------------------------------------
// I(k,a,b) = |sum_j exp(i q_k n(a,b).r_j)|^2 on the quadrature grid.
__global__ __launch_bounds__(256) void k_intensity_grid(const double *__restrict__ atoms, int natoms, int KMAX,
                                                        double dq, const double *__restrict__ gx, int ntheta,
                                                        int nphi, double *__restrict__ I)
{
    __shared__ double sat[3 * 256];
    const int64_t npts = (int64_t)KMAX * ntheta * nphi;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double qx = 0, qy = 0, qz = 0;
    if (i < npts) {
        const int k = (int)(i / ((int64_t)ntheta * nphi)) + 1;
        const int a = (int)((i / nphi) % ntheta), b = (int)(i % nphi);
        const double ct = gx[a], st = sqrt((1.0 - ct) * (1.0 + ct));
        const double ph = 2.0 * 3.14159265358979323846 * b / nphi;
        const double q = k * dq;
        qx = q * st * cos(ph); qy = q * st * sin(ph); qz = q * ct;
    }
    double re = 0.0, im = 0.0;
    for (int j0 = 0; j0 < natoms; j0 += 256) {
        const int nj = min(256, natoms - j0);
        __syncthreads();
        for (int t = threadIdx.x; t < 3 * nj; t += blockDim.x) sat[t] = atoms[3 * j0 + t];
        __syncthreads();
        for (int j = 0; j < nj; ++j) {
            double s, c;
            sincos(qx * sat[3 * j] + qy * sat[3 * j + 1] + qz * sat[3 * j + 2], &s, &c);
            re += c;
            im += s;
        }
    }
    if (i < npts) I[i] = re * re + im * im;
}

// G(k,a,m) = sum_b I(k,a,b) exp(-i m phi_b), m = 0..lfill
__global__ void k_intensity_phi(const double *__restrict__ I, int KMAX, int ntheta, int nphi, int lfill,
                                double2 *__restrict__ G)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)KMAX * ntheta * (lfill + 1);
    if (i >= n) return;
    const int m = (int)(i % (lfill + 1));
    const int64_t ka = i / (lfill + 1);
    const double *row = I + ka * nphi;
    double re = 0, im = 0;
    for (int b = 0; b < nphi; ++b) {
        double s, c;
        sincospi(2.0 * (double)((int64_t)m * b % nphi) / nphi, &s, &c);
        re += row[b] * c;
        im -= row[b] * s;
    }
    G[i] = make_double2(re, im);
}

// c_lm(k) = (2pi/nphi) sum_a w_a sphPlm(l,m,theta_a) G(k,a,m) (m<0 by symmetry)
__global__ void k_intensity_coeff(const double2 *__restrict__ G, const double *__restrict__ P, const double *__restrict__ gw,
                                  int KMAX, int ntheta, int nphi, int lfill, int LMAX,
                                  const int *__restrict__ sean_lm, double2 *__restrict__ coeff)
{
    const int nlm = (lfill + 1) * (lfill + 1);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)KMAX * nlm) return;
    const int k = (int)(i / nlm), q = (int)(i % nlm);
    const int l = (int)sqrt((double)q), m = q - l * l - l;
    const int am = m < 0 ? -m : m;
    double re = 0, im = 0;
    for (int a = 0; a < ntheta; ++a) {
        const double y = P[((int64_t)a * (lfill + 1) + l) * (lfill + 1) + am];
        const double2 g = G[((int64_t)k * ntheta + a) * (lfill + 1) + am];
        re += gw[a] * y * g.x;
        im += gw[a] * y * g.y;
    }
    const double sc = 2.0 * 3.14159265358979323846 / nphi;
    re *= sc; im *= sc;
    if (m < 0) {   // c_{l,-|m|} = (-1)^m conj(c_{l,|m|}) for a real function
        im = -im;
        if (am & 1) { re = -re; im = -im; }
    }
    coeff[(int64_t)k * LMAX * LMAX + sean_lm[q]] = make_double2(re, im);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static int check_ctx(tp_ctx *ctx)
{
    if (!ctx) return tp_fail(TP_E_ARG, "null context");
    TP_HIP(hipSetDevice(ctx->device));
    return 0;
}

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Upload host coefficients of `nshell` shells and pack them to the compact
// layout in dst.
static int upload_coeff(tp_ctx *ctx, tp_basis *b, const double *coeff, int64_t ld, int nshell, double2 *dst,
                        double2 **raw_buf)
{
    if (!coeff || ld < (int64_t)b->LMAX * b->LMAX) return tp_fail(TP_E_ARG, "coefficient vector shorter than LMAX^2");
    double2 *raw = *raw_buf;
    bool tmp = false;
    if (!raw || nshell > b->K || ld != (int64_t)b->LMAX * b->LMAX) {
        if (int rc = dev_alloc(&raw, (size_t)nshell * ld)) return rc;
        tmp = true;
    }
    TP_HIP(hipMemcpyAsync(raw, coeff, sizeof(double2) * (size_t)nshell * ld, hipMemcpyHostToDevice, ctx->stream));
    const int64_t n = (int64_t)nshell * b->n_c;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, raw, ld, nshell,
                       b->d_sean0, b->n_c, dst);
    TP_HIP(hipGetLastError());
    if (tmp) {
        TP_HIP(hipStreamSynchronize(ctx->stream));
        hipFree(raw);
    }
    return 0;
}

// Enqueue one evaluation (As, contraction, final) on coefficients `coef`.
static int enqueue_eval(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, const double2 *coef, double *c3out,
                        double *Eout)
{
    const bool timing = ctx->timing;
    if (timing) TP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    {
        const size_t smem = sizeof(double2) * (2 * (size_t)b->n_c + (size_t)b->max_chunk_groups);
        hipLaunchKernelGGL(k_triple, dim3((unsigned)(b->npairs * b->nchunks)), dim3(256), smem, ctx->stream, coef,
                           b->n_c, b->d_ew, b->d_eidx, b->d_gstart, b->d_ltinfo, b->d_chunks, b->nchunks,
                           b->d_pairs, b->d_As, b->kl_pad);
        TP_HIP(hipGetLastError());
    }
    if (timing) TP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    const int ncols = r ? (int)r->ncols : (int)b->klength;
    const dim3 grid((unsigned)(b->n_rt * b->n_ct));
    if (r && c3out) {
        hipLaunchKernelGGL((k_contract<true, true>), grid, dim3(256), 0, ctx->stream, b->d_F, b->R_pad, b->d_As,
                           b->kl_pad, (int)b->nlt_pad, (int)b->R, ncols, b->n_rt, b->n_ct, r->d_qhat, c3out,
                           b->d_part);
    } else if (r) {
        hipLaunchKernelGGL((k_contract<false, true>), grid, dim3(256), 0, ctx->stream, b->d_F, b->R_pad, b->d_As,
                           b->kl_pad, (int)b->nlt_pad, (int)b->R, ncols, b->n_rt, b->n_ct, r->d_qhat, nullptr,
                           b->d_part);
    } else {
        hipLaunchKernelGGL((k_contract<true, false>), grid, dim3(256), 0, ctx->stream, b->d_F, b->R_pad, b->d_As,
                           b->kl_pad, (int)b->nlt_pad, (int)b->R, ncols, b->n_rt, b->n_ct, nullptr, c3out,
                           nullptr);
    }
    TP_HIP(hipGetLastError());
    if (timing) TP_HIP(hipEventRecord(ctx->ev[2], ctx->stream));
    if (r) {
        hipLaunchKernelGGL(k_final, dim3(1), dim3(1024), 0, ctx->stream, b->d_part, b->n_rt, b->kl_pad, ncols,
                           r->d_qsum, r->d_tf, Eout);
        TP_HIP(hipGetLastError());
    }
    if (timing) TP_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
    return 0;
}

static void record_timing(tp_ctx *ctx)
{
    if (!ctx->timing) return;
    float t;
    for (int i = 0; i < 3; ++i) {
        t = 0;
        if (hipEventElapsedTime(&t, ctx->ev[i], ctx->ev[i + 1]) == hipSuccess) ctx->last_ms[i] = t;
    }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int tp_abi_version(void) { return TP_ABI_VERSION; }

const char *tp_last_error(void) { return g_err.c_str(); }

int tp_init(int device, tp_ctx **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return tp_fail(TP_E_HIP, "no HIP device available");
    if (device < 0 || device >= n) return tp_fail(TP_E_ARG, "device ordinal out of range");
    TP_HIP(hipSetDevice(device));
    tp_ctx *c = new tp_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return tp_fail(TP_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
    for (int i = 0; i < 4; ++i) {
        e = hipEventCreate(&c->ev[i]);
        if (e != hipSuccess) { tp_finalize(c); return tp_fail(TP_E_HIP, "hipEventCreate failed"); }
    }
    *out = c;
    return 0;
}

void tp_finalize(tp_ctx *ctx)
{
    if (!ctx) return;
    hipSetDevice(ctx->device);
    for (int i = 0; i < 4; ++i)
        if (ctx->ev[i]) hipEventDestroy(ctx->ev[i]);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int tp_set_timing(tp_ctx *ctx, int on)
{
    if (!ctx) return tp_fail(TP_E_ARG, "null context");
    ctx->timing = on != 0;
    return 0;
}

int tp_last_timings(const tp_ctx *ctx, double *ms_out, int n)
{
    if (!ctx || !ms_out) return tp_fail(TP_E_ARG, "null argument");
    for (int i = 0; i < n && i < 3; ++i) ms_out[i] = ctx->last_ms[i];
    return 0;
}

void tp_basis_destroy(tp_basis *b)
{
    if (!b) return;
    if (b->ctx) hipSetDevice(b->ctx->device);
    void *ptrs[] = {b->d_F, b->d_ew, b->d_eidx, b->d_gstart, b->d_ltinfo, b->d_chunks, b->d_pairs, b->d_sean0,
                    b->d_As, b->d_part, b->d_out, b->d_raw, b->d_coef, b->d_c3};
    for (void *p : ptrs)
        if (p) hipFree(p);
    delete b;
}

int tp_basis_create(tp_ctx *ctx, int L, int LMAX, int N, int K, double lambda, double dq, int for_intensity,
                    tp_basis **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (int rc = check_ctx(ctx)) return rc;
    if (L < 0 || L > 22 || LMAX < L + 1 || N < 1 || N > 256 || K < 1 || K > 255)
        return tp_fail(TP_E_ARG, "basis parameters out of range (0<=L<=22, LMAX>L, 1<=N<=256, 1<=K<=255)");
    if (lambda != 0.0)
        return tp_fail(TP_E_UNSUPPORTED, "lambda != 0 (curved Ewald sphere) is not implemented in this version");
    tp_basis *b = new tp_basis();
    b->ctx = ctx;
    b->L = L; b->LMAX = LMAX; b->N = N; b->K = K; b->lambda = lambda; b->dq = dq;
    b->for_intensity = for_intensity ? 1 : 0;
    b->lstep = for_intensity ? 2 : 1;
    b->R = 2LL * N * N;
    b->klength = (int64_t)K * (K + 1) * (K + 2) / 6;
    // compact layout
    b->coff.assign(L + 1, -1);
    int nc = 0;
    for (int l = 0; l <= L; l += b->lstep) { b->coff[l] = nc; nc += 2 * l + 1; }
    b->n_c = nc;
    b->sean0.resize(nc);
    for (int l = 0; l <= L; l += b->lstep)
        for (int m = -l; m <= l; ++m) b->sean0[b->coff[l] + m + l] = (int)(tph::seanindex(l, m, LMAX) - 1);

    // ---- enumeration (correlations.jl:92,107-129): l1, l2, l3, m1, m2, m3 with
    // |3j(l1 m1 l2 m2 l3 -m3)| > 100 eps; stored per (lt, mu3) group sorted by m1.
    const double thr = 100.0 * 2.220446049250313e-16;
    struct Ent { double w; int i1, i2; };
    std::vector<double> ew;
    std::vector<uint32_t> eidx;
    std::vector<int> gstart;
    std::vector<int4> ltinfo;
    std::vector<int> lt_entries;     // entries per lt (for chunking)
    std::vector<int> lt_goff;
    // F (lambda = 0): F[r,lt] = sum_{j in lt} w_j Y(l1,m1) Y(l2,m2) Y(l3,mu3) cos(m2 a + mu3 b), Y at theta = pi/2
    const double th = std::acos(0.0);
    std::vector<double> Ytab((size_t)(L + 1) * (2 * L + 1));
    for (int l = 0; l <= L; ++l)
        for (int m = -l; m <= l; ++m) Ytab[(size_t)l * (2 * L + 1) + m + L] = tph::sphPlm(l, m, th);
    auto Y = [&](int l, int m) { return Ytab[(size_t)l * (2 * L + 1) + m + L]; };
    // F accumulates coefficients per (lt, m2, mu3) first: H[lt][(m2,mu3)]
    struct HEnt { int lt, m2, mu3; double h; };
    std::vector<HEnt> hlist;
    int64_t basislen = 0;
    int nlt = 0;
    for (int l1 = 0; l1 <= L; l1 += b->lstep)
        for (int l2 = 0; l2 <= L; l2 += b->lstep)
            for (int l3 = 0; l3 <= L; l3 += b->lstep) {
                if (l3 < std::abs(l1 - l2) || l3 > l1 + l2) continue;
                std::vector<std::vector<Ent>> groups(2 * l3 + 1);
                std::vector<double> hm((size_t)(2 * L + 1) * (2 * L + 1), 0.0);
                int cnt = 0;
                for (int m1 = -l1; m1 <= l1; ++m1)
                    for (int m2 = -l2; m2 <= l2; ++m2) {
                        const int mu3 = -(m1 + m2);
                        if (mu3 < -l3 || mu3 > l3) continue;
                        const double w = tph::wigner3j(l1, m1, l2, m2, l3, mu3);
                        if (!(std::fabs(w) > thr)) continue;
                        groups[mu3 + l3].push_back({w, b->coff[l1] + m1 + l1, b->coff[l2] + m2 + l2});
                        hm[(size_t)(m2 + L) * (2 * L + 1) + (mu3 + L)] += w * Y(l1, m1) * Y(l2, m2) * Y(l3, mu3);
                        ++cnt;
                    }
                if (cnt == 0) continue;
                lt_goff.push_back((int)gstart.size());
                ltinfo.push_back(make_int4(b->coff[l3], (int)gstart.size(), l3, 0));
                for (auto &gv : groups) {
                    gstart.push_back((int)ew.size());
                    for (auto &e : gv) {
                        ew.push_back(e.w);
                        eidx.push_back((uint32_t)e.i1 | ((uint32_t)e.i2 << 16));
                    }
                }
                lt_entries.push_back(cnt);
                for (int m2 = -L; m2 <= L; ++m2)
                    for (int mu3 = -L; mu3 <= L; ++mu3) {
                        const double h = hm[(size_t)(m2 + L) * (2 * L + 1) + (mu3 + L)];
                        if (h != 0.0) hlist.push_back({nlt, m2, mu3, h});
                    }
                basislen += cnt;
                ++nlt;
            }
    gstart.push_back((int)ew.size());
    b->basislen = basislen;
    b->n_lt = nlt;
    b->n_groups = (int)gstart.size() - 1;

    // chunks of l-triples with balanced entry counts, bounded groups (LDS)
    const int64_t target_chunks = std::max<int64_t>(1, std::min<int64_t>(32, basislen / 2048));
    const int64_t per = (basislen + target_chunks - 1) / target_chunks;
    std::vector<int4> chunks;
    {
        int lt0 = 0;
        int64_t acc = 0;
        for (int t = 0; t < nlt; ++t) {
            acc += lt_entries[t];
            const int g0 = lt_goff[lt0], g1 = (t + 1 < nlt) ? lt_goff[t + 1] : b->n_groups;
            if (acc >= per || t == nlt - 1 || g1 - g0 >= 1536) {
                chunks.push_back(make_int4(lt0, t + 1, g0, g1));
                lt0 = t + 1;
                acc = 0;
            }
        }
    }
    b->nchunks = (int)chunks.size();
    b->max_chunk_groups = 0;
    for (auto &c : chunks) b->max_chunk_groups = std::max(b->max_chunk_groups, c.w - c.z);
    std::vector<int> pairs;
    for (int k1 = 1; k1 <= K; ++k1)
        for (int k2 = 1; k2 <= k1; ++k2) pairs.push_back(k1 | (k2 << 16));
    b->npairs = (int)pairs.size();

    // padded dims and F
    b->R_pad = round_up(b->R, CT_BM);
    b->nlt_pad = round_up(nlt, CT_BK);
    b->kl_pad = round_up(b->klength, CT_BN);
    b->n_rt = (int)(b->R_pad / CT_BM);
    b->n_ct = (int)(b->kl_pad / CT_BN);
    std::vector<double> F((size_t)b->nlt_pad * b->R_pad, 0.0);
    {
        std::vector<double> ca((size_t)N * (2 * L + 1)), cb((size_t)2 * N * (2 * L + 1)), sa(ca.size()), sb(cb.size());
        for (int ia = 0; ia < N; ++ia)
            for (int m = -L; m <= L; ++m) {
                const double a = (ia + 0.5) * tph::kPi / N;
                ca[(size_t)ia * (2 * L + 1) + m + L] = std::cos(m * a);
                sa[(size_t)ia * (2 * L + 1) + m + L] = std::sin(m * a);
            }
        for (int jb = 0; jb < 2 * N; ++jb)
            for (int m = -L; m <= L; ++m) {
                const double be = (jb + 0.5) * tph::kPi / N;
                cb[(size_t)jb * (2 * L + 1) + m + L] = std::cos(m * be);
                sb[(size_t)jb * (2 * L + 1) + m + L] = std::sin(m * be);
            }
        for (const HEnt &h : hlist) {
            double *col = F.data() + (size_t)h.lt * b->R_pad;
            for (int jb = 0; jb < 2 * N; ++jb)
                for (int ia = 0; ia < N; ++ia) {
                    // cos(m2 a + mu3 b) = cos(m2 a)cos(mu3 b) - sin(m2 a) sin(mu3 b)  (correlations.jl:141)
                    const double c = ca[(size_t)ia * (2 * L + 1) + h.m2 + L] * cb[(size_t)jb * (2 * L + 1) + h.mu3 + L] -
                                     sa[(size_t)ia * (2 * L + 1) + h.m2 + L] * sb[(size_t)jb * (2 * L + 1) + h.mu3 + L];
                    col[ia + (size_t)jb * N] += h.h * c;
                }
        }
    }

    int rc = 0;
    auto up = [&](auto **dptr, const auto &vec) -> int {
        using T = typename std::remove_reference<decltype(vec)>::type::value_type;
        if (int r2 = dev_alloc(dptr, vec.size())) return r2;
        hipError_t e = hipMemcpy((void *)*dptr, vec.data(), sizeof(T) * vec.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return tp_fail(TP_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return 0;
    };
    if ((rc = up(&b->d_F, F)) || (rc = up(&b->d_ew, ew)) || (rc = up(&b->d_eidx, eidx)) ||
        (rc = up(&b->d_gstart, gstart)) || (rc = up(&b->d_ltinfo, ltinfo)) || (rc = up(&b->d_chunks, chunks)) ||
        (rc = up(&b->d_pairs, pairs)) || (rc = up(&b->d_sean0, b->sean0))) {
        tp_basis_destroy(b);
        return rc;
    }
    if ((rc = dev_alloc(&b->d_As, (size_t)b->nlt_pad * b->kl_pad)) ||
        (rc = dev_alloc(&b->d_part, (size_t)b->n_rt * 2 * b->kl_pad)) || (rc = dev_alloc(&b->d_out, 8)) ||
        (rc = dev_alloc(&b->d_raw, (size_t)K * LMAX * LMAX)) || (rc = dev_alloc(&b->d_coef, (size_t)K * b->n_c))) {
        tp_basis_destroy(b);
        return rc;
    }
    // padding of As (l-triple rows >= n_lt, columns >= klength) must read as zero
    hipError_t e = hipMemset(b->d_As, 0, sizeof(double) * (size_t)b->nlt_pad * b->kl_pad);
    if (e != hipSuccess) { tp_basis_destroy(b); return tp_fail(TP_E_HIP, "hipMemset failed"); }
    *out = b;
    return 0;
}

int tp_basis_info(const tp_basis *b, int64_t *basislen, int64_t *n_lt, int64_t *klength, int64_t *R)
{
    if (!b) return tp_fail(TP_E_ARG, "null basis");
    if (basislen) *basislen = b->basislen;
    if (n_lt) *n_lt = b->n_lt;
    if (klength) *klength = b->klength;
    if (R) *R = b->R;
    return 0;
}

int tp_c3(tp_ctx *ctx, const tp_basis *bc, const double *coeff, int64_t ld, double *c3_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !c3_out) return tp_fail(TP_E_ARG, "null argument");
    tp_basis *b = const_cast<tp_basis *>(bc);
    if (b->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
    if (!b->d_c3)
        if (int rc = dev_alloc(&b->d_c3, (size_t)b->klength * b->R)) return rc;
    if (int rc = upload_coeff(ctx, b, coeff, ld, b->K, b->d_coef, &b->d_raw)) return rc;
    if (int rc = enqueue_eval(ctx, b, nullptr, b->d_coef, b->d_c3, nullptr)) return rc;
    TP_HIP(hipMemcpyAsync(c3_out, b->d_c3, sizeof(double) * (size_t)b->klength * b->R, hipMemcpyDeviceToHost,
                          ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

void tp_ref_destroy(tp_ref *r)
{
    if (!r) return;
    if (r->b && r->b->ctx) hipSetDevice(r->b->ctx->device);
    if (r->d_qhat) hipFree(r->d_qhat);
    if (r->d_qsum) hipFree(r->d_qsum);
    if (r->d_tf) hipFree(r->d_tf);
    delete r;
}

int tp_ref_create(tp_ctx *ctx, const tp_basis *b, const double *c3ref, int Kr, int k3lo, int k3hi, tp_ref **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (int rc = check_ctx(ctx)) return rc;
    if (!b || !c3ref) return tp_fail(TP_E_ARG, "null argument");
    if (b->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
    if (k3lo < 1 || k3hi < k3lo || k3hi > Kr) return tp_fail(TP_E_ARG, "K3_range outside the histogram");
    const int64_t R = b->R, N = b->N;
    // energy(): column index i runs over the K3_range walk (correlations.jl:211-216)
    std::vector<int> k1s, k2s, k3s;
    for (int k1 = k3lo; k1 <= k3hi; ++k1)
        for (int k2 = k3lo; k2 <= k1; ++k2)
            for (int k3 = k3lo; k3 <= k2; ++k3) { k1s.push_back(k1); k2s.push_back(k2); k3s.push_back(k3); }
    const int64_t ncols = (int64_t)k1s.size();
    if (ncols > b->klength) return tp_fail(TP_E_ARG, "K3_range walk longer than the basis' k-triple list");
    std::vector<double> qhat((size_t)ncols * R), qsum(ncols), tf(ncols);
    for (int64_t i = 0; i < ncols; ++i) {
        const int k1 = k1s[i], k2 = k2s[i], k3 = k3s[i];
        const double *q = c3ref + R * ((k3 - 1) + (int64_t)Kr * ((k2 - 1) + (int64_t)Kr * (k1 - 1)));
        double s = 0.0;
        for (int64_t r = 0; r < R; ++r) s += std::fabs(q[r]);
        double t = 0.0;
        for (int64_t r = 0; r < R; ++r) { qhat[(size_t)i * R + r] = q[r] / s; t += q[r] / s; }
        qsum[i] = t;
        // tripletFactor (datagen.jl:166-174)
        tf[i] = (k1 == k2 && k2 == k3) ? 1.0 : ((k1 == k2 || k2 == k3 || k1 == k3) ? 3.0 : 6.0);
    }
    (void)N;
    tp_ref *r = new tp_ref();
    r->b = b;
    r->ncols = ncols;
    if (dev_alloc(&r->d_qhat, qhat.size()) || dev_alloc(&r->d_qsum, ncols) || dev_alloc(&r->d_tf, ncols)) {
        tp_ref_destroy(r);
        return TP_E_OOM;
    }
    if (hipMemcpy(r->d_qhat, qhat.data(), sizeof(double) * qhat.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r->d_qsum, qsum.data(), sizeof(double) * ncols, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(r->d_tf, tf.data(), sizeof(double) * ncols, hipMemcpyHostToDevice) != hipSuccess) {
        tp_ref_destroy(r);
        return tp_fail(TP_E_HIP, "histogram upload failed");
    }
    *out = r;
    return 0;
}

int tp_energy(tp_ctx *ctx, const tp_basis *bc, const tp_ref *r, const double *coeff, int64_t ld, int measure,
              double *E_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !r || !E_out) return tp_fail(TP_E_ARG, "null argument");
    if (measure != 0) { *E_out = 0.0; return 0; }   // correlations.jl:210
    tp_basis *b = const_cast<tp_basis *>(bc);
    if (b->ctx != ctx || r->b != b) return tp_fail(TP_E_STATE, "objects belong to another context/basis");
    if (int rc = upload_coeff(ctx, b, coeff, ld, b->K, b->d_coef, &b->d_raw)) return rc;
    if (int rc = enqueue_eval(ctx, b, r, b->d_coef, nullptr, b->d_out)) return rc;
    TP_HIP(hipMemcpyAsync(E_out, b->d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    record_timing(ctx);
    return 0;
}

void tp_state_destroy(tp_state *s)
{
    if (!s) return;
    if (s->b && s->b->ctx) hipSetDevice(s->b->ctx->device);
    void *ptrs[] = {s->d_cur, s->d_prop, s->d_scal, s->d_step, s->d_flag};
    for (void *p : ptrs)
        if (p) hipFree(p);
    delete s;
}

int tp_state_create(tp_ctx *ctx, const tp_basis *bc, const double *coeff, int64_t ld, int nshell, tp_state **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc) return tp_fail(TP_E_ARG, "null basis");
    tp_basis *b = const_cast<tp_basis *>(bc);
    if (b->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
    if (nshell < b->K) return tp_fail(TP_E_ARG, "state needs at least K shells");
    tp_state *s = new tp_state();
    s->b = b;
    s->nshell = nshell;
    int rc;
    if ((rc = dev_alloc(&s->d_cur, (size_t)nshell * b->n_c)) || (rc = dev_alloc(&s->d_prop, (size_t)nshell * b->n_c)) ||
        (rc = dev_alloc(&s->d_scal, 8)) || (rc = dev_alloc(&s->d_step, std::max(1, b->L / 2))) ||
        (rc = dev_alloc(&s->d_flag, 1))) {
        tp_state_destroy(s);
        return rc;
    }
    double2 *raw = nullptr;
    if ((rc = upload_coeff(ctx, b, coeff, ld, nshell, s->d_cur, &raw))) {
        tp_state_destroy(s);
        return rc;
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) { tp_state_destroy(s); return tp_fail(TP_E_HIP, "state upload failed"); }
    *out = s;
    return 0;
}

int tp_state_read(tp_ctx *ctx, const tp_state *s, double *coeff, int64_t ld)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!s || !coeff) return tp_fail(TP_E_ARG, "null argument");
    const tp_basis *b = s->b;
    if (ld < (int64_t)b->LMAX * b->LMAX) return tp_fail(TP_E_ARG, "ld shorter than LMAX^2");
    std::vector<double2> h((size_t)s->nshell * b->n_c);
    TP_HIP(hipMemcpyAsync(h.data(), s->d_cur, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < s->nshell; ++k)
        for (int c = 0; c < b->n_c; ++c) {
            double *d = coeff + 2 * ((int64_t)k * ld + b->sean0[c]);
            d[0] = h[(size_t)k * b->n_c + c].x;
            d[1] = h[(size_t)k * b->n_c + c].y;
        }
    return 0;
}

int tp_state_energy(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, double *E_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!b || !r || !s || !E_out) return tp_fail(TP_E_ARG, "null argument");
    if (b->ctx != ctx || r->b != b || s->b != b) return tp_fail(TP_E_STATE, "objects belong to another context/basis");
    if (int rc = enqueue_eval(ctx, b, r, s->d_cur, nullptr, b->d_out)) return rc;
    TP_HIP(hipMemcpyAsync(E_out, b->d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    record_timing(ctx);
    return 0;
}

int tp_state_c3(tp_ctx *ctx, const tp_basis *bc, tp_state *s, double *c3_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !s || !c3_out) return tp_fail(TP_E_ARG, "null argument");
    tp_basis *b = const_cast<tp_basis *>(bc);
    if (b->ctx != ctx || s->b != b) return tp_fail(TP_E_STATE, "objects belong to another context/basis");
    if (!b->d_c3)
        if (int rc = dev_alloc(&b->d_c3, (size_t)b->klength * b->R)) return rc;
    if (int rc = enqueue_eval(ctx, b, nullptr, s->d_cur, b->d_c3, nullptr)) return rc;
    TP_HIP(hipMemcpyAsync(c3_out, b->d_c3, sizeof(double) * (size_t)b->klength * b->R, hipMemcpyDeviceToHost,
                          ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tp_anneal(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, const tp_anneal_params *p,
              double *E_trace, double *Enew_trace, int32_t *acc_trace, double *scalars)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!b || !r || !s || !p) return tp_fail(TP_E_ARG, "null argument");
    if (b->ctx != ctx || r->b != b || s->b != b) return tp_fail(TP_E_STATE, "objects belong to another context/basis");
    if (p->nsteps < 0 || p->ntemp < 0 || !(p->stepsizefactor > 0) || !(p->temperature_decay > 0))
        return tp_fail(TP_E_ARG, "invalid annealing parameters");
    if (b->L < 2) return tp_fail(TP_E_ARG, "annealing needs L >= 2");
    if (b->L > 18) return tp_fail(TP_E_UNSUPPORTED, "annealing moves support L <= 18 (2l+1 <= 37)");
    if (!b->for_intensity) return tp_fail(TP_E_UNSUPPORTED, "annealing moves act on even l only");
    const int nl = b->L / 2;
    hipStream_t st = ctx->stream;
    std::vector<double> steps(nl);
    for (int l = 2; l <= b->L; l += 2) steps[l / 2 - 1] = p->initial_stepsize * (l - 1);   // determination.jl:295
    TP_HIP(hipMemcpyAsync(s->d_step, steps.data(), sizeof(double) * nl, hipMemcpyHostToDevice, st));
    int *d_coff = nullptr;
    double *d_tr = nullptr;
    int *d_acc = nullptr;
    const int ns = std::max(1, p->nsteps), nt = std::max(1, p->ntemp);
    int rc = 0;
    if ((rc = dev_alloc(&d_coff, b->coff.size())) || (rc = dev_alloc(&d_tr, 2 * (size_t)ns + nt + 4)) ||
        (rc = dev_alloc(&d_acc, ns))) {
        if (d_coff) hipFree(d_coff);
        if (d_tr) hipFree(d_tr);
        return rc;
    }
    auto cleanup = [&]() { hipFree(d_coff); hipFree(d_tr); hipFree(d_acc); };
    double *d_trE = d_tr, *d_trEn = d_tr + ns, *d_temp = d_tr + 2 * ns;
    TP_HIP(hipMemcpyAsync(d_coff, b->coff.data(), sizeof(int) * b->coff.size(), hipMemcpyHostToDevice, st));
    const int64_t ncoef = (int64_t)s->nshell * b->n_c;
    // initial energy (determination.jl:298) -> scal[0]
    if ((rc = enqueue_eval(ctx, b, r, s->d_cur, nullptr, s->d_scal))) { cleanup(); return rc; }
    // temperature from `ntemp` proposals (calculate_temperature, :244-261)
    double T0 = 0.0;
    if (p->initial_temperature_factor > 0.0 && p->ntemp > 1) {
        for (int i = 0; i < p->ntemp; ++i) {
            hipLaunchKernelGGL(k_propose, dim3(nl), dim3(64), 0, st, s->d_cur, s->d_prop, s->nshell, b->n_c, d_coff,
                               s->d_step, p->seed, (uint32_t)i, TP_TAG_TEMP);
            if ((rc = enqueue_eval(ctx, b, r, s->d_prop, nullptr, b->d_out))) { cleanup(); return rc; }
            hipLaunchKernelGGL(k_store_scalar, dim3(1), dim3(1), 0, st, b->d_out, d_temp, i);
        }
        std::vector<double> et(p->ntemp);
        TP_HIP(hipMemcpyAsync(et.data(), d_temp, sizeof(double) * p->ntemp, hipMemcpyDeviceToHost, st));
        TP_HIP(hipStreamSynchronize(st));
        double m = 0;
        for (double v : et) m += v;
        m /= p->ntemp;
        double v2 = 0;
        for (double v : et) v2 += (v - m) * (v - m);
        T0 = std::sqrt(v2 / (p->ntemp - 1)) * p->initial_temperature_factor;
    }
    TP_HIP(hipMemcpyAsync(s->d_scal + 1, &T0, sizeof(double), hipMemcpyHostToDevice, st));
    double E0 = 0;
    TP_HIP(hipMemcpyAsync(&E0, s->d_scal, sizeof(double), hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    for (int i = 1; i <= p->nsteps; ++i) {
        hipLaunchKernelGGL(k_propose, dim3(nl), dim3(64), 0, st, s->d_cur, s->d_prop, s->nshell, b->n_c, d_coff,
                           s->d_step, p->seed, (uint32_t)i, TP_TAG_MOVE);
        if ((rc = enqueue_eval(ctx, b, r, s->d_prop, nullptr, b->d_out))) { cleanup(); return rc; }
        hipLaunchKernelGGL(k_metropolis, dim3(1), dim3(1), 0, st, s->d_scal, s->d_step, nl, b->d_out, p->seed,
                           (uint32_t)i, p->temperature_decay, p->stepsizefactor, s->d_flag, d_trE, d_trEn, d_acc);
        hipLaunchKernelGGL(k_accept_copy, dim3((unsigned)std::min<int64_t>(64, (ncoef + 255) / 256)), dim3(256), 0,
                           st, s->d_flag, s->d_prop, s->d_cur, ncoef);
    }
    TP_HIP(hipGetLastError());
    double sc[2] = {0, 0};
    TP_HIP(hipMemcpyAsync(sc, s->d_scal, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    if (E_trace && p->nsteps > 0)
        TP_HIP(hipMemcpyAsync(E_trace, d_trE, sizeof(double) * p->nsteps, hipMemcpyDeviceToHost, st));
    if (Enew_trace && p->nsteps > 0)
        TP_HIP(hipMemcpyAsync(Enew_trace, d_trEn, sizeof(double) * p->nsteps, hipMemcpyDeviceToHost, st));
    if (acc_trace && p->nsteps > 0)
        TP_HIP(hipMemcpyAsync(acc_trace, d_acc, sizeof(int) * p->nsteps, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    if (scalars) { scalars[0] = E0; scalars[1] = T0; scalars[2] = sc[0]; scalars[3] = sc[1]; }
    cleanup();
    return 0;
}

int tp_intensity_from_atoms(tp_ctx *ctx, const double *atoms_xyz, int natoms, int KMAX, double dq, int LMAX,
                            int lfill, int ntheta, int nphi, double *coeff_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!atoms_xyz || !coeff_out || natoms < 1 || KMAX < 1 || LMAX < 1 || lfill < 0 || lfill >= LMAX ||
        ntheta < 1 || nphi < 1 || lfill > 22)
        return tp_fail(TP_E_ARG, "invalid intensity parameters");
    hipStream_t st = ctx->stream;
    std::vector<double> gx, gw;
    tph::gauss_legendre(ntheta, gx, gw);
    const int nl1 = lfill + 1;
    std::vector<double> P((size_t)ntheta * nl1 * nl1, 0.0);
    for (int a = 0; a < ntheta; ++a)
        for (int l = 0; l <= lfill; ++l)
            for (int m = 0; m <= l; ++m) P[((size_t)a * nl1 + l) * nl1 + m] = tph::sphPlm(l, m, std::acos(gx[a]));
    std::vector<int> sean((size_t)nl1 * nl1);
    for (int l = 0; l <= lfill; ++l)
        for (int m = -l; m <= l; ++m) sean[l * l + l + m] = (int)(tph::seanindex(l, m, LMAX) - 1);
    double *d_at = nullptr, *d_gx = nullptr, *d_gw = nullptr, *d_I = nullptr, *d_P = nullptr;
    double2 *d_G = nullptr, *d_c = nullptr;
    int *d_sean = nullptr;
    const int64_t npts = (int64_t)KMAX * ntheta * nphi;
    int rc = 0;
    auto cleanup = [&]() {
        void *ps[] = {d_at, d_gx, d_gw, d_I, d_P, d_G, d_c, d_sean};
        for (void *q : ps)
            if (q) hipFree(q);
    };
    if ((rc = dev_alloc(&d_at, 3 * (size_t)natoms)) || (rc = dev_alloc(&d_gx, ntheta)) || (rc = dev_alloc(&d_gw, ntheta)) ||
        (rc = dev_alloc(&d_I, npts)) || (rc = dev_alloc(&d_P, P.size())) ||
        (rc = dev_alloc(&d_G, (size_t)KMAX * ntheta * nl1)) || (rc = dev_alloc(&d_c, (size_t)KMAX * LMAX * LMAX)) ||
        (rc = dev_alloc(&d_sean, sean.size()))) {
        cleanup();
        return rc;
    }
    hipMemcpyAsync(d_at, atoms_xyz, sizeof(double) * 3 * natoms, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d_gx, gx.data(), sizeof(double) * ntheta, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d_gw, gw.data(), sizeof(double) * ntheta, hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d_P, P.data(), sizeof(double) * P.size(), hipMemcpyHostToDevice, st);
    hipMemcpyAsync(d_sean, sean.data(), sizeof(int) * sean.size(), hipMemcpyHostToDevice, st);
    hipMemsetAsync(d_c, 0, sizeof(double2) * (size_t)KMAX * LMAX * LMAX, st);
    hipLaunchKernelGGL(k_intensity_grid, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, d_at, natoms, KMAX, dq,
                       d_gx, ntheta, nphi, d_I);
    const int64_t ng = (int64_t)KMAX * ntheta * nl1;
    hipLaunchKernelGGL(k_intensity_phi, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, d_I, KMAX, ntheta, nphi,
                       lfill, d_G);
    const int64_t nc = (int64_t)KMAX * nl1 * nl1;
    hipLaunchKernelGGL(k_intensity_coeff, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, d_G, d_P, d_gw, KMAX,
                       ntheta, nphi, lfill, LMAX, d_sean, d_c);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(coeff_out, d_c, sizeof(double2) * (size_t)KMAX * LMAX * LMAX, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return tp_fail(TP_E_HIP, std::string("intensity kernels: ") + hipGetErrorString(e));
    return 0;
}

}  // extern "C"

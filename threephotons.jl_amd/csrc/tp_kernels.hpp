// tp_kernels.hpp -- gfx950 device code of libtp (included by tp_hip.hip).
//
// One c3 + Bayes-energy evaluation (lambda = 0) is four launches:
//   k_triple_mfma  coupling-form triple products As[lt][kt]                 (FP64 MFMA, LDS)
//                  (k_triple: the VALU form for K > 27)
//   k_gemm         G  = H . As      (folded (m2,mu3) x k-triple)            (FP64 MFMA)
//   k_c3sep        c3 = Bm . G on the separated (alpha, beta) grid, fused
//                  Bayes partial sums (k_gemm with the energy epilogue for N > 32)
//   k_final        per-column energies, fixed-order total, optional Metropolis step
// H[(m2,mu3), lt] = sum_{j in lt} w_j Y(l1,m1) Y(l2,m2) Y(l3,mu3) and
// Bm[r, (m2,mu3)] = cos(m2 a_r + mu3 b_r), so Bm . H is the l-triple basis F
// of the reference's slow path (flab, correlations.jl:337-362) and
// Bm . H . As = B . PA (correlations.jl:182).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "tp_fastlog.hpp"

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2 __attribute__((ext_vector_type(2)));   // promotable to registers, unlike HIP's double2

#define TP_TAG_MOVE 0x4D4F5645u
#define TP_TAG_ACC 0x41434350u
#define TP_TAG_TEMP 0x54454D50u

constexpr double kPiD = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller normals.  Streams are
// keyed by the run seed; the counter is (j>>1, l-index, step, tag).
// ---------------------------------------------------------------------------
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

__device__ inline double dev_normal(uint64_t seed, uint32_t step, uint32_t li, uint32_t j, uint32_t tag)
{
    uint32_t c[4] = {j >> 1, li, step, tag};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double a = u01(c[0], c[1]), b = u01(c[2], c[3]);
    const double r = sqrt(-2.0 * log(1.0 - a));
    const double ang = 2.0 * kPiD * b;
    return (j & 1) ? r * sin(ang) : r * cos(ang);
}

__device__ inline double dev_uniform(uint64_t seed, uint32_t step, uint32_t tag)
{
    uint32_t c[4] = {0u, 0xFFFFFFFFu, step, tag};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return u01(c[0], c[1]);
}

// Double-buffered coefficients: buffer b1 is used when (*sel ^ flip) is set.
__device__ __forceinline__ const double2 *pick(const double2 *b0, const double2 *b1, const int *sel, int flip)
{
    const int s = (sel ? sel[0] : 0) ^ flip;
    return s ? b1 : b0;
}

// Gather the compact (even l <= L) coefficients out of LMAX^2-strided vectors.
// Zero n doubles (a fill kernel, so that no memset node is needed).
__global__ __launch_bounds__(256) void k_zero64(double *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_pack(const double2 *__restrict__ raw, int64_t ld, int nshell, const int *__restrict__ sean0,
                       int n_c, double2 *__restrict__ coef)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nshell * n_c) return;
    const int k = (int)(i / n_c), c = (int)(i % n_c);
    coef[i] = raw[(int64_t)k * ld + sean0[c]];
}

// ---------------------------------------------------------------------------
// Triple products in coupling form.  Workgroup = one chunk of consecutive
// l-triples (its (lt,mu3) groups g0..g1 and their weights, staged in LDS);
// all K shells' coefficients sit in LDS and the workgroup walks every k-pair
// (k1 >= k2) in batches of PB:
//   stage 1: T[p][lt,mu3] = sum_{m1} w(l1 m1 l2 m2 l3 mu3) c_k1[l1,m1] c_k2[l2,m2],
//            m2 = -m1-mu3 (one thread per (pair, group));
//   stage 2: As[lt][kt(k1,k2,k3)] = Re sum_mu3 T[p][lt,mu3] c_k3[l3,mu3], k3 <= k2
//            (one thread per (pair, lt, k3), k3 fastest: As rows are written
//            contiguously).
// This equals sum_{j in lt} w_j Re(c_k1[i1] c_k2[i2] c_k3[i3])
// (correlations.jl:165-174, cuda_kernel.cu:30-37) regrouped by mu3.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_triple(const double2 *__restrict__ cb0, const double2 *__restrict__ cb1,
                                                 const int *__restrict__ sel, int flip, int n_c, int K,
                                                 const double *__restrict__ ew, const uint32_t *__restrict__ eidx,
                                                 const int *__restrict__ gstart, const int4 *__restrict__ ltinfo,
                                                 const int4 *__restrict__ chunks, const int *__restrict__ pairs,
                                                 int npairs, int PB, int Gmax, int Emax, double *__restrict__ As,
                                                 int64_t kl_pad)
{
    extern __shared__ double2 sm[];
    const double2 *coef = pick(cb0, cb1, sel, flip);
    double2 *sc = sm;                                          // [K][n_c]
    double2 *T = sc + (size_t)K * n_c;                         // [PB][G]
    double *sw = reinterpret_cast<double *>(T + (size_t)PB * Gmax);   // [Emax]
    uint32_t *si = reinterpret_cast<uint32_t *>(sw + Emax);            // [Emax]
    int *sg = reinterpret_cast<int *>(si + Emax);                      // [Gmax + 1]
    const int4 cw = chunks[blockIdx.x];
    const int lt0 = cw.x, lt1 = cw.y, g0 = cw.z, g1 = cw.w;
    const int G = g1 - g0, nlt = lt1 - lt0;
    const int tid = threadIdx.x, bs = blockDim.x;
    const int e0 = gstart[g0], ne = gstart[g1] - e0;
    for (int i = tid; i < K * n_c; i += bs) sc[i] = coef[i];
    for (int i = tid; i < ne; i += bs) {
        sw[i] = ew[e0 + i];
        si[i] = eidx[e0 + i];
    }
    for (int i = tid; i <= G; i += bs) sg[i] = gstart[g0 + i] - e0;
    __syncthreads();
    for (int pb = 0; pb < npairs; pb += PB) {
        const int nb = min(PB, npairs - pb);
        for (int idx = tid; idx < nb * G; idx += bs) {
            const int q = idx / G, gl = idx - q * G;
            const int pk = pairs[pb + q], k1 = pk & 0xffff, k2 = pk >> 16;
            const double2 *c1 = sc + (k1 - 1) * n_c, *c2 = sc + (k2 - 1) * n_c;
            const int s = sg[gl], e = sg[gl + 1];
            double tr0 = 0.0, ti0 = 0.0, tr1 = 0.0, ti1 = 0.0;
            int j = s;
            for (; j + 1 < e; j += 2) {
                const uint32_t ia = si[j], ib = si[j + 1];
                const double wa = sw[j], wb = sw[j + 1];
                const double2 a0 = c1[ia & 0xffffu], b0 = c2[ia >> 16];
                const double2 a1 = c1[ib & 0xffffu], b1 = c2[ib >> 16];
                tr0 += wa * (a0.x * b0.x - a0.y * b0.y);
                ti0 += wa * (a0.x * b0.y + a0.y * b0.x);
                tr1 += wb * (a1.x * b1.x - a1.y * b1.y);
                ti1 += wb * (a1.x * b1.y + a1.y * b1.x);
            }
            if (j < e) {
                const uint32_t ia = si[j];
                const double wa = sw[j];
                const double2 a0 = c1[ia & 0xffffu], b0 = c2[ia >> 16];
                tr0 += wa * (a0.x * b0.x - a0.y * b0.y);
                ti0 += wa * (a0.x * b0.y + a0.y * b0.x);
            }
            T[idx] = make_double2(tr0 + tr1, ti0 + ti1);
        }
        __syncthreads();
        for (int idx = tid; idx < nb * nlt * K; idx += bs) {
            const int k3 = idx % K + 1, rest = idx / K;
            const int q = rest / nlt, lt = lt0 + (rest - q * nlt);
            const int pk = pairs[pb + q], k1 = pk & 0xffff, k2 = pk >> 16;
            if (k3 > k2) continue;
            const int4 li = ltinfo[lt];
            const double2 *Tl = T + q * G + (li.y - g0);
            const double2 *c3 = sc + (k3 - 1) * n_c + li.x;
            const int n = 2 * li.z + 1;
            double s0 = 0.0, s1 = 0.0;
            int m = 0;
            for (; m + 1 < n; m += 2) {
                const double2 t0 = Tl[m], c0 = c3[m], t1 = Tl[m + 1], c1 = c3[m + 1];
                s0 += t0.x * c0.x - t0.y * c0.y;
                s1 += t1.x * c1.x - t1.y * c1.y;
            }
            if (m < n) {
                const double2 t0 = Tl[m], c0 = c3[m];
                s0 += t0.x * c0.x - t0.y * c0.y;
            }
            const int64_t kt = (int64_t)(k1 - 1) * k1 * (k1 + 1) / 6 + (int64_t)(k2 - 1) * k2 / 2 + (k3 - 1);
            As[(int64_t)lt * kl_pad + kt] = s0 + s1;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Triple products on FP64 MFMA (K <= 28).  Workgroup = (l-triple lt, half h
// of the k1 range: k1 in 16h+1..16h+16), 16 waves (the LDS allows one
// workgroup per CU; 16 waves give 4 per SIMD to hide the LDS latency).  The
// l1, l2, l3 blocks of all K shells sit in LDS.  Iteration t handles
// mu3 = 16t..16t+15, one per wave:
//   stage 1: [Tr | Ti](k1, k2) = A'.B' over the group's E entries, where
//            A'[k1][e] = w_e Re c_k1[i1_e], A'[k1][E+e] = w_e Im c_k1[i1_e],
//            B'[e] = [Re c_k2[i2_e] | Im c_k2[i2_e]],
//            B'[E+e] = [-Im c_k2[i2_e] | Re c_k2[i2_e]]
//            (complex T = sum_e a_e b_e as one real GEMM with K-dim 2E);
//            the (k1 >= k2) entries go to Tf[slot][pair] in LDS, slot =
//            (mu3 - 16t, re/im) (slot-major keeps lanes on distinct banks);
//   stage 2: As[pair][k3] += sum_slot Tf[slot][pair] . [Re c_k3[l3,mu3] ; -Im c_k3[l3,mu3]],
//            accumulated across all iterations in MFMA registers.
// For k1 <= 16 the k2 > 16 and k3 > 16 blocks are skipped.  Same sums as
// k_triple below, regrouped for MFMA.
// ---------------------------------------------------------------------------
// Stage-1 MFMA operands of one k-step: lane (row l16, k lq) gets
// A'[k1s][4 ks + lq] and B'[4 ks + lq][k2] for the two k2 blocks (Tr and Ti
// columns).  Zero past 2E or past K.
struct S1Ops {
    double a, br0, bi0, br1, bi1;
};

__device__ __forceinline__ S1Ops s1_fetch(int ks, int E, int s, int K, int k1s, int l16, int lq, int n1, int n2,
                                          const double *sw, const uint32_t *si, const double2 *c1b,
                                          const double2 *c2b, bool two)
{
    S1Ops o = {0.0, 0.0, 0.0, 0.0, 0.0};
    const int ep = 4 * ks + lq;
    if (ep < 2 * E) {
        const bool first = ep < E;
        const int e = first ? ep : ep - E;
        const double w = sw[s + e];
        const uint32_t id = si[s + e];
        const int i1 = id & 0xffff, i2 = id >> 16;
        if (k1s <= K) {
            const double2 c = c1b[(k1s - 1) * n1 + i1];
            o.a = w * (first ? c.x : c.y);
        }
        const int k2a = l16 + 1, k2b = 16 + l16 + 1;
        if (k2a <= K) {
            const double2 c = c2b[(k2a - 1) * n2 + i2];
            o.br0 = first ? c.x : -c.y;
            o.bi0 = first ? c.y : c.x;
        }
        if (two && k2b <= K) {
            const double2 c = c2b[(k2b - 1) * n2 + i2];
            o.br1 = first ? c.x : -c.y;
            o.bi1 = first ? c.y : c.x;
        }
    }
    return o;
}

// Row length of the per-shell l-block norms |c_k,l|^2 (nrm2[k][l/2], even l <= 30).
constexpr int kNrmLd = 16;

// Column of k-triple kt in an As row: kt-tile j = kt / 64 of an evaluation
// sits in 64-column block cbase + j * nprop + p (evaluations of one batch
// group interleaved by tile; a single evaluation has cbase = 0, nprop = 1,
// p = 0, i.e. column kt).
__device__ __forceinline__ int64_t as_col(int64_t kt, int cbase, int nprop, int p)
{
    return ((int64_t)cbase + (kt >> 6) * nprop + p) * 64 + (kt & 63);
}

__device__ __forceinline__ void triple_mfma_body(double2 *sm, const double2 *__restrict__ coef, int n_c, int K, int lt,
                                                 int h, const double *__restrict__ ew,
                                                 const uint32_t *__restrict__ eidx, const int *__restrict__ gstart,
                                                 const int4 *__restrict__ ltinfo, int Lw, int Emax, int Prows,
                                                 double *__restrict__ As, int64_t ldA, int cbase, int nprop, int p0)
{
    const int4 li = ltinfo[lt];
    const int l3 = li.z, l1 = li.w & 0xff, l2 = (li.w >> 8) & 0xff;
    const int n1 = 2 * l1 + 1, n2 = 2 * l2 + 1, n3 = 2 * l3 + 1;
    const int goff = li.y;
    // LDS carve-up
    double2 *c1b = sm;                                   // [K][n1]
    double2 *c2b = c1b + (size_t)K * Lw;                 // [K][n2]
    double2 *c3b = c2b + (size_t)K * Lw;                 // [K][n3]
    double *Tf = reinterpret_cast<double *>(c3b + (size_t)K * Lw);   // [32 slots][Prows]
    double *sw = Tf + 32 * (size_t)Prows;                            // [Emax]
    uint32_t *si = reinterpret_cast<uint32_t *>(sw + Emax);          // [Emax]
    int *sg = reinterpret_cast<int *>(si + Emax);                    // [n3 + 1]
    int *pk = sg + (n3 + 1);                                         // [Prows]

    constexpr int NT = 1024;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int co1 = l1 * (l1 - 1) / 2, co2 = l2 * (l2 - 1) / 2, co3 = l3 * (l3 - 1) / 2;   // compact offsets (even l)
    for (int i = tid; i < K * n1; i += NT) c1b[i] = coef[(size_t)(i / n1) * n_c + co1 + i % n1];
    for (int i = tid; i < K * n2; i += NT) c2b[i] = coef[(size_t)(i / n2) * n_c + co2 + i % n2];
    for (int i = tid; i < K * n3; i += NT) c3b[i] = coef[(size_t)(i / n3) * n_c + co3 + i % n3];
    const int e0 = gstart[goff], ne = gstart[goff + n3] - e0;
    for (int i = tid; i < ne; i += NT) {
        sw[i] = ew[e0 + i];
        const uint32_t id = eidx[e0 + i];
        si[i] = (uint32_t)((int)(id & 0xffffu) - co1) | ((uint32_t)((int)(id >> 16) - co2) << 16);
    }
    for (int i = tid; i <= n3; i += NT) sg[i] = gstart[goff + i] - e0;
    // pairs (k1 >= k2) of this half, in the reference order, rows of Tf
    const int k1lo = 16 * h + 1, k1hi = min(16 * h + 16, K);
    const int off_h = (k1lo - 1) * k1lo / 2;                     // pairs with k1 < k1lo
    const int P = k1hi * (k1hi + 1) / 2 - off_h;
    const int nrb = (P + 15) / 16;
    for (int p = tid; p < nrb * 16; p += NT) {
        if (p < P) {
            const int g = p + off_h;
            int k1 = (int)((1.0 + sqrt(1.0 + 8.0 * g)) * 0.5);
            while (k1 * (k1 - 1) / 2 > g) --k1;
            while ((k1 + 1) * k1 / 2 <= g) ++k1;
            pk[p] = k1 | ((g - k1 * (k1 - 1) / 2 + 1) << 8);
        } else {
            pk[p] = 0;
        }
    }
    for (int i = tid; i < 32 * Prows; i += NT) Tf[i] = 0.0;
    __syncthreads();

    dbl4 acc2[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) acc2[j] = dbl4{0.0, 0.0, 0.0, 0.0};
    const int nob = h ? 2 * nrb : nrb;   // stage-2 output blocks (<= 32)
    const int niter = (n3 + 15) / 16;
    const int k1s = 16 * h + l16 + 1;   // stage-1 A row of this lane
    for (int t = 0; t < niter; ++t) {
        const int mu = 16 * t + wave;    // index mu3 + l3 of this wave's group
        if (mu < n3) {
            const int s = sg[mu], E = sg[mu + 1] - s;
            const int nks = (2 * E + 3) / 4;
            dbl4 a1[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) a1[j] = dbl4{0.0, 0.0, 0.0, 0.0};
            // operands of k-step ks are fetched from LDS one step ahead of the MFMAs
            S1Ops cur = s1_fetch(0, E, s, K, k1s, l16, lq, n1, n2, sw, si, c1b, c2b, h > 0);
            for (int ks = 0; ks < nks; ++ks) {
                S1Ops nxt = cur;
                if (ks + 1 < nks) nxt = s1_fetch(ks + 1, E, s, K, k1s, l16, lq, n1, n2, sw, si, c1b, c2b, h > 0);
                a1[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a, cur.br0, a1[0], 0, 0, 0);
                a1[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a, cur.bi0, a1[2], 0, 0, 0);
                if (h > 0) {   // k1 <= 16 needs no k2 > 16
                    a1[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a, cur.br1, a1[1], 0, 0, 0);
                    a1[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(cur.a, cur.bi1, a1[3], 0, 0, 0);
                }
                cur = nxt;
            }
            // rows k1 = 16h + lq + 4v + 1, columns k2 = 16*(j&1) + l16 + 1
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if ((j & 1) && h == 0) continue;
                const int k2 = 16 * (j & 1) + l16 + 1;
                const int slot = 2 * wave + (j >> 1);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k1 = 16 * h + lq + 4 * v + 1;
                    if (k1 <= K && k2 <= k1) {
                        const int p = (k1 - 1) * k1 / 2 + (k2 - 1) - off_h;
                        Tf[slot * Prows + p] = a1[j][v];
                    }
                }
            }
        }
        __syncthreads();
        // stage 2: blocks b = wave + 16j -> (row block b>>1, k3 block b&1); for
        // h = 0 only k3 block 0 (k3 <= k2 <= 16): b -> (b, 0).  k-steps kk
        // cover mu3 offsets 2kk, 2kk+1: only those below n3 are run
        const int kkn = (min(16, n3 - 16 * t) + 1) >> 1;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int b = wave + 16 * j;
            if (b < nob) {
                const int rb = h ? b >> 1 : b, cb = h ? b & 1 : 0;
                const int k3 = 16 * cb + l16 + 1;
#pragma unroll
                for (int kk = 0; kk < 8; ++kk) {
                    if (kk >= kkn) break;   // uniform over the workgroup
                    const int slot = 4 * kk + lq;
                    const double av = Tf[slot * Prows + 16 * rb + l16];
                    const int m3 = 16 * t + (slot >> 1);
                    double bv = 0.0;
                    if (m3 < n3 && k3 <= K) {
                        const double2 c = c3b[(k3 - 1) * n3 + m3];
                        bv = (slot & 1) ? -c.y : c.x;
                    }
                    acc2[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc2[j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // rows = pairs p = 16 rb + lq + 4v, columns k3 = 16 cb + l16 + 1
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int b = wave + 16 * j;
        if (b < nob) {
            const int rb = h ? b >> 1 : b, cb = h ? b & 1 : 0;
            const int k3 = 16 * cb + l16 + 1;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int p = 16 * rb + lq + 4 * v;
                if (p < P) {
                    const int k1 = pk[p] & 0xff, k2 = pk[p] >> 8;
                    if (k3 <= k2) {
                        const int64_t kt = (int64_t)(k1 - 1) * k1 * (k1 + 1) / 6 + (int64_t)(k2 - 1) * k2 / 2 + (k3 - 1);
                        As[(int64_t)lt * ldA + as_col(kt, cbase, nprop, p0)] = acc2[j][v];
                    }
                }
            }
        }
    }
}

__global__ __launch_bounds__(1024) void k_triple_mfma(const double2 *__restrict__ cb0, const double2 *__restrict__ cb1,
                                                     const int *__restrict__ sel, int flip, int n_c, int K,
                                                     const double *__restrict__ ew, const uint32_t *__restrict__ eidx,
                                                     const int *__restrict__ gstart, const int4 *__restrict__ ltinfo,
                                                     const int *__restrict__ work, int Lw, int Emax, int Prows,
                                                     double *__restrict__ As, int64_t kl_pad)
{
    extern __shared__ double2 sm[];
    const int wk = work[blockIdx.x];
    triple_mfma_body(sm, pick(cb0, cb1, sel, flip), n_c, K, wk & 0xffff, wk >> 16, ew, eidx, gstart, ltinfo, Lw, Emax,
                     Prows, As, kl_pad, 0, 1, 0);
}

// ---------------------------------------------------------------------------
// G = H . As of one evaluation in FP64 (the class-weight contraction of the
// reference's B . PA, correlations.jl:179-193, before the separated grid):
// G[class][kt] = sum_lt H[lt][class] As[lt][kt], H K-major (ld ldh, rows
// >= hrows read as zero), K = kdim l-triples (a multiple of 16, padding rows
// zero).  Workgroup = 16 k-triple columns x 16 RB classes; its 4 waves take
// a quarter of K each, so they share no operand: the main loop stages
// nothing through LDS and has no barrier, each wave's RB accumulators are
// independent MFMA chains, and its operand loads (1 As + RB H values per lane
// and k-step, from L2) run PF k-steps ahead.  The four partial tiles are
// summed in LDS in wave order (deterministic).  Workgroups of one column
// block are dispatched to one XCD (block id mod 8), whose L2 then serves its
// As columns to all row tiles.
// ---------------------------------------------------------------------------
template <int RB, int PF>
__global__ __launch_bounds__(256) void k_dgemm_g(const double *__restrict__ H, int64_t ldh, int hrows,
                                                 const double *__restrict__ As, int64_t ldA, int kdim, int ncb,
                                                 int nrt, double *__restrict__ G, int64_t ldg, int grows)
{
    __shared__ double red[4][RB][4][64];   // [wave][row block][v][lane]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int cpx = (ncb + 7) >> 3, idx = blockIdx.x >> 3;
    const int cb = (blockIdx.x & 7) * cpx + idx / nrt, rt = idx % nrt;
    if (cb >= ncb) return;   // whole workgroup, before the barrier
    const int c0 = cb * 16, r0 = rt * 16 * RB;
    const int kq = kdim >> 2, nks = kq >> 2;   // l-triples and k-steps of this wave
    const double *pb = As + (int64_t)(wave * kq + lq) * ldA + c0 + l16;
    // H rows at or past hrows: loaded from the last row (in range) and zeroed
    // where used, without a branch
    const double *pa[RB];
    bool rok[RB];
#pragma unroll
    for (int t = 0; t < RB; ++t) {
        const int row = r0 + 16 * t + l16;
        rok[t] = row < hrows;
        pa[t] = H + (int64_t)(wave * kq + lq) * ldh + min(row, hrows - 1);
    }
    double qa[PF][RB], qb[PF];
    auto load = [&](int slot, int ks) {
        const int64_t o = (int64_t)(4 * ks);
        qb[slot] = pb[o * ldA];
#pragma unroll
        for (int t = 0; t < RB; ++t) qa[slot][t] = pa[t][o * ldh];   // masked where used, so never waited on here
    };
    // slot p holds k-step k0 + p; prefetch indices are clamped to the last
    // k-step (in range; the extra loads go unused)
#pragma unroll
    for (int p = 0; p < PF; ++p) load(p, min(p, nks - 1));
    dbl4 acc[RB];
#pragma unroll
    for (int t = 0; t < RB; ++t) acc[t] = dbl4{0.0, 0.0, 0.0, 0.0};
    // k-steps past nks (the last group of PF) run with zero operands: no
    // remainder branch, which would make the compiler move the accumulators
    // between register files in every k-step
    for (int k0 = 0; k0 < nks; k0 += PF) {
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            const bool kok = k0 + p < nks;
            const double b = qb[p];
            double a[RB];
#pragma unroll
            for (int t = 0; t < RB; ++t) a[t] = rok[t] && kok ? qa[p][t] : 0.0;
            load(p, min(k0 + p + PF, nks - 1));
#pragma unroll
            for (int t = 0; t < RB; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b, acc[t], 0, 0, 0);
        }
    }
    // partial tiles: rows 16 t + lq + 4 v, column l16
#pragma unroll
    for (int t = 0; t < RB; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[wave][t][v][lane] = acc[t][v];
    __syncthreads();
    for (int i = tid; i < RB * 256; i += 256) {
        const int t = i >> 8, v = (i >> 6) & 3, ln = i & 63;
        const double g = ((red[0][t][v][ln] + red[1][t][v][ln]) + red[2][t][v][ln]) + red[3][t][v][ln];
        const int row = r0 + 16 * t + (ln >> 4) + 4 * v;
        if (row < grows) G[(int64_t)row * ldg + c0 + (ln & 15)] = g;
    }
}

// ---------------------------------------------------------------------------
// FP64 MFMA GEMM  C[M x N] = A[M x K] . B[K x N]  (v_mfma_f64_16x16x4f64).
// A is stored K-major (A[k*lda + m]), B as B[k*ldb + n]; K is a multiple of
// 16 and the padded operands hold zeros.  A BM x BN tile per workgroup of 4
// waves arranged WGM x (4/WGM); K staged through LDS 16 at a time with the
// next slice prefetched into registers while the MFMAs of the current one
// run.  Workgroups are mapped so that each XCD (block id mod 8) owns
// consecutive row tiles, whose A rows then stay in that XCD's L2.
//   OUT = 1: C written column-major out[col*ldo + r] for r < M, col < ncols
//            (the c3raw layout of return_raw, correlations.jl:182)
//   OUT = 2: the full padded tile written row-major out[r*ldo + col]
//   ENERGY : per column, the partial sums over this tile's rows of
//            p = max(c,1e-30) and qhat*log(p) (correlations.jl:203-220) go to
//            part[(rt*2+{0,1})*ldp + col]; requires WGM = 4.
// ---------------------------------------------------------------------------
template <int BM, int BN, int WGM, int OUT, bool ENERGY, bool XCDC = false>
__global__ __launch_bounds__(256) void k_gemm(const double *__restrict__ A, int64_t lda,
                                              const double *__restrict__ B, int64_t ldb, int Kdim, int M, int ncols,
                                              int n_rt, int n_ct, double *__restrict__ out, int64_t ldo,
                                              const double *__restrict__ qhat, double *__restrict__ part, int64_t ldp,
                                              const double *__restrict__ logtab)
{
    __shared__ double s_lt[ENERGY ? kLogTabSize : 1];
    constexpr int BK = 16;
    constexpr int WGN = 4 / WGM;
    constexpr int WTM = BM / WGM, WTN = BN / WGN;
    constexpr int RB = WTM / 16, CB = WTN / 16;
    constexpr int A2 = BK * BM / 2 / 256;
    constexpr int B2 = BK * BN / 2 / 256;
    static_assert(A2 >= 1 && B2 >= 1 && RB >= 1 && CB >= 1, "tile too small");
    static_assert(!ENERGY || WGM == 4, "energy epilogue needs all waves on the same columns");
    __shared__ double sA[BK][BM];
    __shared__ double sB[BK][BN];
    __shared__ double red[ENERGY ? 4 : 1][2][BN];

    const int nblk = n_rt * n_ct;
    int rt, ct;
    if (XCDC) {
        // XCD x (= block id mod 8) owns column tiles x cpx .. x cpx + cpx - 1
        // with all their row tiles, consecutive in its dispatch order, so each
        // B column block is fetched from HBM once, into that XCD's L2; the
        // grid is 8 cpx n_rt (blocks past n_ct exit)
        const int cpx = (n_ct + 7) >> 3, i = blockIdx.x >> 3;
        ct = (blockIdx.x & 7) * cpx + i / n_rt;
        rt = i % n_rt;
        if (ct >= n_ct) return;   // whole workgroup, before any barrier
    } else {
        int lin = blockIdx.x;
        if ((nblk & 7) == 0) lin = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
        rt = lin / n_ct;
        ct = lin % n_ct;
    }
    if (ENERGY)
        for (int i = threadIdx.x; i < kLogTabSize; i += blockDim.x) s_lt[i] = logtab[i];
    __syncthreads();
    const int r0 = rt * BM, c0 = ct * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % WGM, wn = wave / WGM;
    const int l16 = lane & 15, lq = lane >> 4;

    dbl4 acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

    dbl2 ra[A2], rb[B2];
#pragma unroll
    for (int q = 0; q < A2; ++q) {
        const int i = tid + 256 * q, row = i / (BM / 2), c2 = i % (BM / 2);
        ra[q] = *reinterpret_cast<const dbl2 *>(A + (int64_t)row * lda + r0 + 2 * c2);
    }
#pragma unroll
    for (int q = 0; q < B2; ++q) {
        const int i = tid + 256 * q, row = i / (BN / 2), c2 = i % (BN / 2);
        rb[q] = *reinterpret_cast<const dbl2 *>(B + (int64_t)row * ldb + c0 + 2 * c2);
    }
    for (int k0 = 0; k0 < Kdim; k0 += BK) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < A2; ++q) {
            const int i = tid + 256 * q, row = i / (BM / 2), c2 = i % (BM / 2);
            *reinterpret_cast<dbl2 *>(&sA[row][2 * c2]) = ra[q];
        }
#pragma unroll
        for (int q = 0; q < B2; ++q) {
            const int i = tid + 256 * q, row = i / (BN / 2), c2 = i % (BN / 2);
            *reinterpret_cast<dbl2 *>(&sB[row][2 * c2]) = rb[q];
        }
        __syncthreads();
        if (k0 + BK < Kdim) {
#pragma unroll
            for (int q = 0; q < A2; ++q) {
                const int i = tid + 256 * q, row = i / (BM / 2), c2 = i % (BM / 2);
                ra[q] = *reinterpret_cast<const dbl2 *>(A + (int64_t)(k0 + BK + row) * lda + r0 + 2 * c2);
            }
#pragma unroll
            for (int q = 0; q < B2; ++q) {
                const int i = tid + 256 * q, row = i / (BN / 2), c2 = i % (BN / 2);
                rb[q] = *reinterpret_cast<const dbl2 *>(B + (int64_t)(k0 + BK + row) * ldb + c0 + 2 * c2);
            }
        }
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            double a[RB], b[CB];
#pragma unroll
            for (int i = 0; i < RB; ++i) a[i] = sA[kk * 4 + lq][wm * WTM + i * 16 + l16];
#pragma unroll
            for (int j = 0; j < CB; ++j) b[j] = sB[kk * 4 + lq][wn * WTN + j * 16 + l16];
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
        const int col = c0 + wn * WTN + j * 16 + l16;
        const bool cok = col < ncols;
        double sp = 0.0, sq = 0.0;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const int rbase = r0 + wm * WTM + i * 16 + lq;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rbase + 4 * v;
                const double c = acc[i][j][v];
                if (OUT == 2) out[(int64_t)r * ldo + col] = c;
                if (cok && r < M) {
                    if (OUT == 1) out[(int64_t)col * ldo + r] = c;
                    if (ENERGY) {
                        const double pv = c > 1e-30 ? c : 1e-30;
                        const double q = qhat[(int64_t)col * M + r];
                        sp += pv;
                        sq += q * tp_fastlog(pv, s_lt);
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
            part[((int64_t)rt * 2 + 0) * ldp + c0 + tid] = sp;
            part[((int64_t)rt * 2 + 1) * ldp + c0 + tid] = sq;
        }
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision contraction c3 = F . As (the l-triple basis F = Bm . H,
// precomputed in FP64 and split on the host).
//   TP precision 1 (FP16 x3): x = hi + lo with hi = fp16(x), lo = fp16(x - hi);
//     F.As ~ Fhi.Ahi + (Fhi.Alo + Flo.Ahi), the main product and the two
//     correction products in separate FP32 accumulators, summed in FP64.
//   TP precision 2 (BF16): one BF16 product, FP32 accumulation.
// Operands are scaled by powers of two (exact) to sit in [2^13, 2^14): F by
// one global factor, As per k-triple column (k_split_as).
// ---------------------------------------------------------------------------
typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 b16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint16_t f16_bits(double x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
__device__ __forceinline__ double f16_val(uint16_t b) { return (double)__builtin_bit_cast(_Float16, b); }
__device__ __forceinline__ uint16_t bf16_bits(double x) { return __builtin_bit_cast(uint16_t, (__bf16)(float)x); }
// x = hi + lo + e with |e| <= 2^-22 |x| (x scaled into the FP16 range): x is
// rounded to FP32 first and split there with the hardware conversions (the
// direct double -> half conversion is a long software sequence on gfx950);
// xf - hi is exact in FP32.
__device__ __forceinline__ void f16_split(double x, uint16_t &hb, uint16_t &lb)
{
    float xf = (float)x;
    // xf is materialised, so that the compiler cannot fold the two roundings
    // into one FP64 -> FP16 conversion (a long integer sequence on gfx950):
    // the conversions stay v_cvt_f32_f64 and v_cvt_f16_f32
    __asm__("" : "+v"(xf));
    const _Float16 h = (_Float16)xf;
    hb = __builtin_bit_cast(uint16_t, h);
    lb = __builtin_bit_cast(uint16_t, (_Float16)(xf - (float)h));
}

// As [lt][kt] (FP64) -> 16-bit copies in [kt][lt] layout (lt contiguous, row
// length ldk), one 64 x 64 tile per block, with a per-column power-of-two
// scale s_kt (stored as 1/s) from the bound
//   |As(kt, lt)| <= |c_k1,l1| |c_k2,l2| |c_k3,l3| <= n_k1 n_k2 n_k3,
// n_k = max_l |c_k,l| (the 3j coupling of three unit vectors has modulus <= 1
// by the orthogonality of the Clebsch-Gordan map), so |As s| < 2^14.
// The shell norms n_k come from k_shell_scales (nk[k]), computed once per
// evaluation.
template <int MODE>
__global__ __launch_bounds__(256) void k_split_as(const double *__restrict__ As, int64_t kl_pad, int nlt, int ldk,
                                                  const double *__restrict__ nk, int K, int klength,
                                                  uint16_t *__restrict__ Ahi, uint16_t *__restrict__ Alo,
                                                  double *__restrict__ inv_scale)
{
    __shared__ double tile[64][65];
    __shared__ double s_nk[64];
    __shared__ double s_sc[64];
    const int kt0 = blockIdx.x * 64, lt0 = blockIdx.y * 64;
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
    if (tid < K) s_nk[tid] = nk[tid];
    __syncthreads();
    if (tid < 64) {
        double s = 1.0;
        const int kt = kt0 + tid;
        if (kt < klength) {
            // kt -> (k1, k2, k3) in the order k1 >= k2 >= k3 (correlations.jl:72-82)
            int k1 = 1;
            while ((int64_t)(k1 + 1) * (k1 + 2) * k1 / 6 <= kt) ++k1;
            const int r1 = kt - (k1 - 1) * k1 * (k1 + 1) / 6;
            int k2 = 1;
            while ((k2 + 1) * k2 / 2 <= r1) ++k2;
            const int k3 = r1 - (k2 - 1) * k2 / 2 + 1;
            const double bnd = s_nk[k1 - 1] * s_nk[k2 - 1] * s_nk[k3 - 1];
            if (bnd > 0.0) s = ldexp(1.0, 13 - ilogb(bnd));
        }
        s_sc[tid] = s;
        if (blockIdx.y == 0) inv_scale[kt] = 1.0 / s;
    }
    const int kr = tid >> 2, q = tid & 3;   // write phase: row kt0+kr, 16 consecutive lt
    {
        for (int r = g; r < 64; r += 4) {
            const int lt = lt0 + r;
            tile[r][c] = lt < nlt ? As[(int64_t)lt * kl_pad + kt0 + c] : 0.0;
        }
        __syncthreads();
        const double s = s_sc[kr];
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
        u16x8 *ph = reinterpret_cast<u16x8 *>(Ahi + (int64_t)(kt0 + kr) * ldk + lt0 + q * 16);
        ph[0] = h0;
        ph[1] = h1;
        if (MODE == 1) {
            u16x8 *pl = reinterpret_cast<u16x8 *>(Alo + (int64_t)(kt0 + kr) * ldk + lt0 + q * 16);
            pl[0] = l0;
            pl[1] = l1;
        }
        __syncthreads();
    }
}

// C[r][kt] = sum_k A[r][k] B[kt][k] (both operands k-contiguous, 16-bit),
// BM x BN tile (BM, BN in {64, 128}) per step of a workgroup of 4 waves
// (2 x 2), K staged through LDS 32 at a time (rows padded to 40 halves:
// conflict-free 16-byte reads), next slice prefetched into registers.  MFMA
// v_mfma_f32_16x16x32_{f16,bf16}: lane holds A[l%16][8(l/16)..+7],
// B[8(l/16)..+7][l%16], accumulator rows 4(l/16)+v (measured on gfx950).
// FP16 hi/lo: hi.hi and (hi.lo + lo.hi) in two FP32 accumulators, summed in
// FP64 (the corrections then round at 2^-11 of the main sum's scale).
// Epilogue as k_gemm: raw c3 (OUT = 1) and/or per-column partial energy sums.
// BATCH > 0: the columns are 64-column blocks of many evaluations; block cb
// reads its histogram columns from qtab[cb] (nvalid[cb] of them are real).
//   BATCH = 1: one workgroup per (row tile, column tile), column-tile-major
//              (the row tiles of a column tile share an XCD's L2);
//   BATCH = 2: one workgroup per column tile walking all row tiles (its B tile
//              read once; the energy sums it writes, part row 0, are whole
//              columns);
//   BATCH = 3: as 1, with the row tiles partitioned over the 8 XCDs.
// Either way the column tiles that share histogram columns are neighbours in
// the (XCD-grouped) tile order.
constexpr int H_BM = 128, H_BN = 64, H_BK = 32, H_LDS_ROW = 40;

template <int BM, int BN, int MODE, int OUT, bool ENERGY, int BATCH>
__global__ __launch_bounds__(256) void k_gemm16(const uint16_t *__restrict__ Ahi, const uint16_t *__restrict__ Alo,
                                                int64_t lda, const uint16_t *__restrict__ Bhi,
                                                const uint16_t *__restrict__ Blo, int64_t ldb, int Kdim,
                                                double a_inv, const double *__restrict__ b_inv, int M, int ncols,
                                                int n_rt, int n_ct, double *__restrict__ out, int64_t ldo,
                                                const double *__restrict__ qhat, double *__restrict__ part,
                                                int64_t ldp, const double *const *__restrict__ qtab,
                                                const int *__restrict__ nvalid)
{
    constexpr int NP = MODE == 1 ? 2 : 1;   // hi (+ lo) planes
    constexpr int WTM = BM / 2, WTN = BN / 2, RB = WTM / 16, CB = WTN / 16;
    constexpr int QA = BM * 4 / 256, QB = BN * 4 / 256;   // 16-byte chunks per thread per plane
    __shared__ uint16_t sA[NP][BM][H_LDS_ROW];
    __shared__ uint16_t sB[NP][BN][H_LDS_ROW];
    __shared__ double red[2][2][BN];

    const int nblk = BATCH == 2 ? n_ct : n_rt * n_ct;
    int lin = blockIdx.x;
    if ((nblk & 7) == 0) lin = (blockIdx.x & 7) * (nblk >> 3) + (blockIdx.x >> 3);
    int ct, rt_first;
    if (BATCH == 3) {
        // the XCD of workgroup id (id mod 8) owns row tiles [x n_rt/8, (x+1) n_rt/8)
        // for every column tile: its L2 keeps those rows of A; the column tiles
        // of one k-triple tile (all proposals) run back to back and share the
        // histogram tile (host: n_rt % 8 == 0 and nblk % 8 == 0)
        const int rpx = n_rt >> 3, x = blockIdx.x & 7, i = blockIdx.x >> 3;
        rt_first = x * rpx + i % rpx;
        ct = i / rpx;
    } else {
        ct = BATCH == 2 ? lin : (BATCH == 1 ? lin / n_rt : lin % n_ct);
        rt_first = BATCH == 2 ? 0 : (BATCH == 1 ? lin % n_rt : lin / n_ct);
    }
    const int rt_last = BATCH == 2 ? n_rt : rt_first + 1;
    const int c0 = ct * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int l16 = lane & 15, lq = lane >> 4;
    auto gl = [](const uint16_t *base, int64_t ld, int row, int k) {
        return *reinterpret_cast<const u16x8 *>(base + (int64_t)row * ld + k);
    };

    // per-column data of this lane's CB columns
    bool cok[CB];
    const double *qcol[CB];
    double sc[CB], sp[CB], sq[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) {
        const int col = c0 + wn * WTN + j * 16 + l16;
        qcol[j] = nullptr;
        if (BATCH) {
            cok[j] = (col & 63) < nvalid[col >> 6];
            if (ENERGY && cok[j]) qcol[j] = qtab[col >> 6] + (int64_t)(col & 63) * M;
        } else {
            cok[j] = col < ncols;
            if (ENERGY) qcol[j] = qhat + (int64_t)col * M;
        }
        sc[j] = cok[j] ? a_inv * b_inv[col] : 0.0;
        sp[j] = 0.0;
        sq[j] = 0.0;
    }

    for (int rt = rt_first; rt < rt_last; ++rt) {
        const int r0 = rt * BM;
        f32x4 acc[RB][CB], cor[RB][CB];
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int j = 0; j < CB; ++j) {
                acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
                cor[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
            }
        u16x8 pa[NP][QA], pb[NP][QB];
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int i = tid + 256 * q, row = i >> 2, kc = (i & 3) * 8;
            pa[0][q] = gl(Ahi, lda, r0 + row, kc);
            if (NP == 2) pa[NP - 1][q] = gl(Alo, lda, r0 + row, kc);
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int i = tid + 256 * q, row = i >> 2, kc = (i & 3) * 8;
            pb[0][q] = gl(Bhi, ldb, c0 + row, kc);
            if (NP == 2) pb[NP - 1][q] = gl(Blo, ldb, c0 + row, kc);
        }
        for (int k0 = 0; k0 < Kdim; k0 += H_BK) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < QA; ++q) {
                const int i = tid + 256 * q, row = i >> 2, kc = (i & 3) * 8;
#pragma unroll
                for (int p = 0; p < NP; ++p) *reinterpret_cast<u16x8 *>(&sA[p][row][kc]) = pa[p][q];
            }
#pragma unroll
            for (int q = 0; q < QB; ++q) {
                const int i = tid + 256 * q, row = i >> 2, kc = (i & 3) * 8;
#pragma unroll
                for (int p = 0; p < NP; ++p) *reinterpret_cast<u16x8 *>(&sB[p][row][kc]) = pb[p][q];
            }
            __syncthreads();
            if (k0 + H_BK < Kdim) {
#pragma unroll
                for (int q = 0; q < QA; ++q) {
                    const int i = tid + 256 * q, row = i >> 2, kc = k0 + H_BK + (i & 3) * 8;
                    pa[0][q] = gl(Ahi, lda, r0 + row, kc);
                    if (NP == 2) pa[NP - 1][q] = gl(Alo, lda, r0 + row, kc);
                }
#pragma unroll
                for (int q = 0; q < QB; ++q) {
                    const int i = tid + 256 * q, row = i >> 2, kc = k0 + H_BK + (i & 3) * 8;
                    pb[0][q] = gl(Bhi, ldb, c0 + row, kc);
                    if (NP == 2) pb[NP - 1][q] = gl(Blo, ldb, c0 + row, kc);
                }
            }
            u16x8 fa[NP][RB], fb[NP][CB];
#pragma unroll
            for (int p = 0; p < NP; ++p) {
#pragma unroll
                for (int i = 0; i < RB; ++i)
                    fa[p][i] = *reinterpret_cast<const u16x8 *>(&sA[p][wm * WTM + i * 16 + l16][8 * lq]);
#pragma unroll
                for (int j = 0; j < CB; ++j)
                    fb[p][j] = *reinterpret_cast<const u16x8 *>(&sB[p][wn * WTN + j * 16 + l16][8 * lq]);
            }
#pragma unroll
            for (int i = 0; i < RB; ++i)
#pragma unroll
                for (int j = 0; j < CB; ++j) {
                    if (MODE == 1) {
                        const h16x8 ah = __builtin_bit_cast(h16x8, fa[0][i]), bh = __builtin_bit_cast(h16x8, fb[0][j]);
                        const h16x8 al = __builtin_bit_cast(h16x8, fa[NP - 1][i]);
                        const h16x8 bl = __builtin_bit_cast(h16x8, fb[NP - 1][j]);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[i][j], 0, 0, 0);
                        cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, cor[i][j], 0, 0, 0);
                        cor[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, cor[i][j], 0, 0, 0);
                    } else {
                        const b16x8 ah = __builtin_bit_cast(b16x8, fa[0][i]), bh = __builtin_bit_cast(b16x8, fb[0][j]);
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[i][j], 0, 0, 0);
                    }
                }
        }
        // epilogue of this row tile
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const int col = c0 + wn * WTN + j * 16 + l16;
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                const int rbase = r0 + wm * WTM + i * 16 + 4 * lq;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = rbase + v;
                    if (cok[j] && r < M) {
                        const double c = ((double)acc[i][j][v] + (MODE == 1 ? (double)cor[i][j][v] : 0.0)) * sc[j];
                        if (OUT == 1) out[(int64_t)col * ldo + r] = c;
                        if (ENERGY) {
                            const double pv = c > 1e-30 ? c : 1e-30;
                            sp[j] += pv;
                            sq[j] += qcol[j][r] * tp_log_series(pv);   // table-free: no LDS for the log
                        }
                    }
                }
            }
        }
    }
    if (ENERGY) {
        const int rt_out = BATCH == 2 ? 0 : rt_first;
#pragma unroll
        for (int j = 0; j < CB; ++j) {
            const int colt = wn * WTN + j * 16 + l16;
            double a = sp[j], b = sq[j];
            a += __shfl_xor(a, 16);
            b += __shfl_xor(b, 16);
            a += __shfl_xor(a, 32);
            b += __shfl_xor(b, 32);
            if (lq == 0) {
                red[wm][0][colt] = a;
                red[wm][1][colt] = b;
            }
        }
        __syncthreads();
        if (tid < BN) {
            part[((int64_t)rt_out * 2 + 0) * ldp + c0 + tid] = red[0][0][tid] + red[1][0][tid];
            part[((int64_t)rt_out * 2 + 1) * ldp + c0 + tid] = red[0][1][tid] + red[1][1][tid];
        }
    }
}

// ---------------------------------------------------------------------------
// Energy finish and the annealing step.
//   E_kt = tF * (qsum * log(S) - Q),  S = sum_rt part0, Q = sum_rt part1,
// in one workgroup of kFinalThreads: thread t sums the columns t, t +
// kFinalThreads, ..., then a fixed tree, so E is bitwise reproducible.  One
// workgroup needs no cross-workgroup ordering: with one per 32 columns and
// the last to arrive summing the block totals, the agent-scope release
// fence each block needs (an L2 write-back on each XCD) made the kernel 8 us
// at 3,276 columns, against 3 us.
// mode 0: E -> Eout;  mode 1: also E -> temp[step] (temperature probe);
// mode 2: also scal[0] = E (initial energy);
// mode 3: Metropolis step (evaluate_step, determination.jl:491-515, with
//         T *= decay first, rotate_all_at_once :328): on acceptance the step
//         sizes grow by `factor`, E is replaced and the proposal buffer
//         becomes current (*sel ^= 1); otherwise step sizes shrink.  With a
//         control block (ctl != nullptr) a step is skipped whole once ctl[0]
//         is set; after each step ctl[1] = its index, ctl[2] counts the
//         accepted steps, and ctl[0] is set when the step size of the largest
//         l is at or below minstep (the loop condition of rotate_all_at_once,
//         determination.jl:1,322).  Trace entry of step i: i - 1 - tr0.
// ---------------------------------------------------------------------------
struct McArgs {
    int mode;
    uint32_t step;
    uint64_t seed;
    double decay, factor;
    int nl;
    double *scal;        // {E, T}
    double *stepsizes;   // [nl]
    int *sel;            // current buffer index
    double *trE, *trEn;  // [steps of this call]
    int *trA;            // [steps of this call]
    double *temp;        // [ntemp]
    uint32_t tr0;        // step index before the first step of this call
    double minstep;      // stopping rule (<= 0: none)
    int *ctl;            // {stopped, last step, accepted steps} or nullptr
    uint64_t *seqp;      // mode 0: after E, seq is stored here with system-scope release (host polls it)
    uint64_t seq;
};

constexpr int kFinalThreads = 1024;   // k_final's one workgroup

// What follows an evaluation's energy E (one thread): the energy out, and in
// the Monte Carlo modes the temperature probe (1), the initial energy (2) or
// the Metropolis decision of step mc.step (3; evaluate_step,
// determination.jl:491-515).
__device__ __forceinline__ void final_energy(double E, double *__restrict__ Eout, McArgs mc)
{
    Eout[0] = E;
    if (mc.seqp) __hip_atomic_store(mc.seqp, mc.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    if (mc.mode == 1) {
        mc.temp[mc.step] = E;
    } else if (mc.mode == 2) {
        mc.scal[0] = E;
    } else if (mc.mode == 3) {
        if (mc.ctl && mc.ctl[0]) return;   // the run has met its stopping rule
        const double T = mc.scal[1] * mc.decay;
        mc.scal[1] = T;
        const double dE = E - mc.scal[0];
        const double ap = exp(-dE / T);
        const double u = dev_uniform(mc.seed, mc.step, TP_TAG_ACC);
        const int acc = (dE < 0.0 || (dE > 0.0 && u < ap)) ? 1 : 0;
        if (acc) {
            for (int i = 0; i < mc.nl; ++i) mc.stepsizes[i] *= mc.factor;
            mc.scal[0] = E;
            mc.sel[0] ^= 1;
        } else {
            for (int i = 0; i < mc.nl; ++i) mc.stepsizes[i] /= mc.factor;
        }
        const uint32_t t = mc.step - 1 - mc.tr0;
        mc.trE[t] = mc.scal[0];
        mc.trEn[t] = E;
        mc.trA[t] = acc;
        if (mc.ctl) {
            mc.ctl[1] = (int)mc.step;
            mc.ctl[2] += acc;
            if (mc.minstep > 0.0 && mc.stepsizes[mc.nl - 1] <= mc.minstep) mc.ctl[0] = 1;
        }
    }
}

__global__ __launch_bounds__(kFinalThreads) void k_final(const double *__restrict__ part, int n_rt, int64_t ldp,
                                                         int ncols, const double *__restrict__ qsum,
                                                         const double *__restrict__ tf, double *__restrict__ Eout,
                                                         McArgs mc)
{
    __shared__ double fin[kFinalThreads / 64];
    const int tid = threadIdx.x;
    double e = 0.0;
    for (int kt = tid; kt < ncols; kt += kFinalThreads) {
        double S = 0.0, Q = 0.0;
#pragma unroll 4
        for (int q = 0; q < n_rt; ++q) {
            S += part[((int64_t)q * 2 + 0) * ldp + kt];
            Q += part[((int64_t)q * 2 + 1) * ldp + kt];
        }
        e += tf[kt] * (qsum[kt] * log(S) - Q);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
    if ((tid & 63) == 0) fin[tid >> 6] = e;
    __syncthreads();
    if (tid != 0) return;
    double E = 0.0;
#pragma unroll
    for (int w = 0; w < kFinalThreads / 64; ++w) E += fin[w];
    final_energy(E, Eout, mc);
}

// ---------------------------------------------------------------------------
// One Monte Carlo proposal (get_MC_step, determination.jl:232-241): for each
// even l in 2..L (one workgroup per l) a rotation by beta = stepsize_l * z in
// the plane of a random orthonormal 2-frame (u,v) of R^(2l+1),
//   R = I + (cos b - 1)(u u^T + v v^T) + sin b (v u^T - u v^T),
// which is inv(C) G(beta) C of random_rotation_step (utilities.jl:180-190)
// written through the two rows of C that G touches.  It acts on the complex
// vectors through the real basis, x <- ctr^H R ctr x, with ctr = Umat(l)
// (spherical_harmonics.jl:445-457) in its two-entries-per-row form.  Lane i
// of a wave owns entry m = i - l of one shell; the four waves share the
// shells of the workgroup's chunk (blockIdx.y: 4 shells per workgroup, so
// the launch has L/2 x nshell/4 workgroups; each recomputes its l's frame
// from the same Philox counters).  The current buffer is read, the other one
// written; the l = 2 workgroups also copy the l = 0 block.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(256) void k_propose(double2 *__restrict__ b0, double2 *__restrict__ b1,
                                                 const int *__restrict__ sel, int nshell, int n_c,
                                                 const int *__restrict__ coff_l,
                                                 const double *__restrict__ stepsizes, uint64_t seed,
                                                 uint32_t step, uint32_t tag, const int *__restrict__ stopped)
{
    __shared__ double uv[2][64];
    __shared__ double sbeta;
    if (stopped && stopped[0]) return;   // no step follows a met stopping rule
    const int s = sel ? sel[0] : 0;
    const double2 *cur = s ? b1 : b0;
    double2 *prop = s ? b0 : b1;
    const int li = blockIdx.x, l = 2 * (li + 1), n = 2 * l + 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (wave == 0) {
        // the frame: Gram-Schmidt of two Gaussian vectors, lane i = entry i
        const double g1 = lane < n ? dev_normal(seed, step, (uint32_t)li, (uint32_t)lane, tag) : 0.0;
        const double g2 = lane < n ? dev_normal(seed, step, (uint32_t)li, (uint32_t)(n + lane), tag) : 0.0;
        const double u = g1 / sqrt(wave_sum(g1 * g1));
        const double d = wave_sum(u * g2);
        const double vt = g2 - d * u;
        const double v = vt / sqrt(wave_sum(vt * vt));
        uv[0][lane] = u;
        uv[1][lane] = v;
        if (lane == 0) sbeta = stepsizes[li] * dev_normal(seed, step, (uint32_t)li, (uint32_t)(2 * n), tag);
    }
    __syncthreads();
    const double cm = cos(sbeta) - 1.0, sb = sin(sbeta);
    const double h = 0.70710678118654752440;
    const int off = coff_l[l];
    const double u = lane < n ? uv[0][lane] : 0.0, v = lane < n ? uv[1][lane] : 0.0;
    const int m = lane - l;                      // this lane's order m (valid when lane < n)
    const int am = m < 0 ? -m : m;
    const double sg = (am & 1) ? -1.0 : 1.0;
    for (int k = 4 * blockIdx.y + wave; k < nshell && k < 4 * blockIdx.y + 4; k += 4) {
        if (li == 0 && lane == 0) prop[(int64_t)k * n_c + coff_l[0]] = cur[(int64_t)k * n_c + coff_l[0]];
        const double2 *x = cur + (int64_t)k * n_c + off;
        double2 *xo = prop + (int64_t)k * n_c + off;
        double2 xs = make_double2(0.0, 0.0), xm = make_double2(0.0, 0.0);
        if (lane < n) {
            xs = x[lane];            // x_m
            xm = x[2 * l - lane];    // x_{-m}
        }
        // y = ctr x:  y_0 = x_0; y_u = (x_u + (-1)^u x_{-u})/sqrt2 (u > 0);
        //             y_{-u} = (i (-1)^u x_{-u} - i x_u)/sqrt2 (u > 0)
        double yr, yi;
        if (m == 0) {
            yr = xs.x; yi = xs.y;
        } else if (m > 0) {
            yr = h * (xs.x + sg * xm.x); yi = h * (xs.y + sg * xm.y);
        } else {   // here xs = x_{-u}, xm = x_u with u = -m
            yr = h * (-sg * xs.y + xm.y); yi = h * (sg * xs.x - xm.x);
        }
        if (lane >= n) { yr = 0.0; yi = 0.0; }
        const double ar = wave_sum(u * yr), ai = wave_sum(u * yi);
        const double br = wave_sum(v * yr), bi = wave_sum(v * yi);
        yr += cm * (u * ar + v * br) + sb * (v * ar - u * br);
        yi += cm * (u * ai + v * bi) + sb * (v * ai - u * bi);
        // x = ctr^H y:  x_m = (y_m + i y_{-m})/sqrt2 ; x_{-m} = (-1)^m (y_m - i y_{-m})/sqrt2
        const int mir = (lane < n) ? (2 * l - lane) : lane;
        const double zr = __shfl(yr, mir), zi = __shfl(yi, mir);   // y at -m
        if (lane < n) {
            double orr, oi;
            if (m == 0) {
                orr = yr; oi = yi;
            } else if (m > 0) {
                orr = h * (yr - zi); oi = h * (yi + zr);
            } else {   // own y is y_{-|m|}, z is y_{|m|}
                orr = h * sg * (zr + yi); oi = h * sg * (zi - yr);
            }
            xo[lane] = make_double2(orr, oi);
        }
    }
}

// ---------------------------------------------------------------------------
// Synthetic intensity of point atoms (input generator, not the hot path).
// ---------------------------------------------------------------------------
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
        const double ph = 2.0 * kPiD * b / nphi;
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
__global__ void k_intensity_coeff(const double2 *__restrict__ G, const double *__restrict__ P,
                                  const double *__restrict__ gw, int KMAX, int ntheta, int nphi, int lfill, int LMAX,
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
    const double sc = 2.0 * kPiD / nphi;
    re *= sc;
    im *= sc;
    if (m < 0) {   // c_{l,-|m|} = (-1)^m conj(c_{l,|m|}) for a real function
        im = -im;
        if (am & 1) { re = -re; im = -im; }
    }
    coeff[(int64_t)k * LMAX * LMAX + sean_lm[q]] = make_double2(re, im);
}

// ---------------------------------------------------------------------------
// c3 = Bm . G by separation of the (alpha, beta) grid, FP64 MFMA.
// Bm[r,(m2,mu3)] = cos(m2 a_i + mu3 b_j) = cos(m2 a_i) cos(mu3 b_j) - sin(m2 a_i) sin(mu3 b_j)
// with r = i + N j, so per k-triple column
//   stage A: C[m2][j] = sum_{mu3} cos(mu3 b_j) G[(m2,mu3)],  S[m2][j] = sum_{mu3} sin(mu3 b_j) G[(m2,mu3)]
//   stage B: c3[i][j] = sum_{m2} cos(m2 a_i) C[m2][j] - sin(m2 a_i) S[m2][j]
// (8 n_m N + 4 n_m2 R flop per column instead of 2 R n_m; L = 18, N = 32:
// 1.17e5 instead of 5.57e5).  Workgroup
// = (16 columns, 16 values of j); classes of one m2 are contiguous
// (m2start).  Tables: TA[cs][class][j] = cos / sin(mu3_class b_j) (ld jpad),
// TB[i][2 m2i + cs] = cos(m2 a_i) / -sin(m2 a_i) (ld kb).  Energy partial
// sums of the workgroup's rows go to part row blockIdx.y (as k_gemm does per
// row tile); OUT = 1 also writes c3 (column-major, ld ldo).
//
// FOLD (energy only, even N, every class with even m2 and even mu3, i.e.
// lambda dq = 0): on the grids a_i = (i + 1/2) pi / N, b_j = (j + 1/2) pi / N
// (alpharange, alpharange_2pi; correlations.jl:247-256) c3 takes the same
// value on the four bins (i, j), (i, j + N), (N-1-i, 2N-1-j), (N-1-i, N-1-j):
// b + pi changes mu3 b by a multiple of 2 pi, and (pi - a, 2 pi - b) turns
// m2 a + mu3 b into a multiple of 2 pi minus itself.  The four bins of every
// orbit are distinct, and the orbits of i < N/2, j < N cover the grid once,
// so the kernel forms c3 on those N^2/2 bins only: S gets 4 p per bin and Q
// the folded histogram's sum of the orbit's four q values (qhat [column]
// [i + (N/2) j]), the same sums as over the whole grid.  A quarter of stage
// B, of the logarithms and of the histogram reads, half of stage A.
// ---------------------------------------------------------------------------
constexpr int kSepMaxKB = 24;   // 2 (L/2 + 1) padded to 4, L <= 22
// The wide instance (QB = kSepWideQB) takes the odd-m classes of the curved
// Ewald sphere: 2 (L + 1) rows of (m2, cos/sin) padded to 4, L <= 19, and
// any number of classes per m2 (stage A runs them in chunks of 24).  Its LDS
// image (87 KB at L = 18) leaves one workgroup per CU.
constexpr int kSepWideQB = 10;
constexpr int kSepWideMaxKB = 4 * kSepWideQB;

// Dynamic LDS of k_c3sep: CS[kb][16][17] doubles.
inline size_t c3sep_lds_bytes(int kb) { return sizeof(double) * (size_t)kb * 16 * 17; }

template <int OUT, bool ENERGY, bool BATCH = false, bool FOLD = false, int QB = 6>
__global__ __launch_bounds__(512, QB > 6 ? 1 : 3) void k_c3sep(const double *__restrict__ G, int64_t ldg, const int *__restrict__ m2start,
                                               int n_m2, const double *__restrict__ TA, int64_t ta_plane, int jpad,
                                               const double *__restrict__ TB, int kb, int N, int ncols,
                                               const double *__restrict__ qhat, double *__restrict__ part,
                                               int64_t ldp, double *__restrict__ out, int64_t ldo,
                                               const double *__restrict__ logtab,
                                               const double *const *__restrict__ qtab = nullptr,
                                               const int *__restrict__ nvalid = nullptr)
{
    static_assert(!FOLD || (ENERGY && OUT == 0), "the folded grid forms the energy only");
    static_assert(QB == 6 || (QB == kSepWideQB && !FOLD), "wide instance: the full grid");
    __shared__ __align__(16) double s_lt[ENERGY ? kLogTabSize : 1];
    __shared__ double red[8][2][16];
    extern __shared__ double cs_dyn[];
    // CS[k][j local][col] = cs_dyn[(k * 16 + j) * 17 + col] (padded row)
    auto CS = [&](int k, int r, int c) -> double & { return cs_dyn[(k * 16 + r) * 17 + c]; };
    if (ENERGY)   // ordered before the epilogue by the barrier after stage A
        for (int i = threadIdx.x; i < kLogTabSize; i += 512) s_lt[i] = logtab[i];
    const LogConsts lk = tp_fastlog_consts(logtab);   // uniform: scalar loads
    // 8 waves.  The histogram values (HBM) are requested after stage A's
    // operands: loads complete in order, so issuing them first would make
    // stage A wait for them.  Three workgroups fit a CU (dynamic LDS sized kb).
    constexpr int NW = 8, MAXB = 4;   // stage-B blocks per wave: 16 j x (N/16) i-blocks / 8 waves (N <= 32)
    const int c0 = blockIdx.x * 16, jc = blockIdx.y, j0 = jc * 16;
    // bins: i < NI, j < NJ; histogram column of RQ values, bin (i, j) at i + NI j
    const int NI = FOLD ? N / 2 : N, NJ = FOLD ? N : 2 * N, RQ = NI * NJ;
    // BATCH: the 16 columns lie in one 64-column block of a batch, whose
    // histogram columns start at qtab[block] (nvalid[block] of them are real)
    const int nv = BATCH ? nvalid[c0 >> 6] - (c0 & 63) : ncols - c0;
    const double *qbase = nullptr;
    if (ENERGY) {
        if (!BATCH) qbase = qhat + (int64_t)c0 * RQ;
        else if (nv > 0) qbase = qtab[c0 >> 6] + (int64_t)(c0 & 63) * RQ;
        else qbase = qtab[c0 >> 6];   // no real column: loads from the block's base, results unused
    }
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform: scalar table loads
    const int l16 = lane & 15, lq = lane >> 4;
    const int nib = (NI + 15) / 16, nob = 16 * nib;
    // padding rows of CS (k' >= 2 n_m2) stay zero
    for (int i = tid; i < (kb - 2 * n_m2) * 256; i += 512) {
        const int k = 2 * n_m2 + i / 256, r = (i / 16) % 16, c = i % 16;
        CS(k, r, c) = 0.0;
    }
    // ---- stage A: blocks (m2i, cs), rows j local, K = classes of m2 (<= 24;
    // the wide instance: any number, in chunks of 24)
    for (int b = wave; b < 2 * n_m2; b += NW) {
        const int m2i = b >> 1, cs = b & 1;
        const int cb0 = m2start[m2i], ce = m2start[m2i + 1];
        const double *ta = TA + (int64_t)cs * ta_plane;
        dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};
        for (int cb = cb0; cb < ce; cb += 24) {
            double av[6], gv[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) {
                const int cls = cb + 4 * q + lq;
                const bool ok = cls < ce;
                av[q] = ok ? ta[(int64_t)cls * jpad + j0 + l16] : 0.0;
                gv[q] = ok ? G[(int64_t)cls * ldg + c0 + l16] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 6; ++q)
                if (cb + 4 * q < ce) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[q], gv[q], acc, 0, 0, 0);
            if constexpr (QB == 6) break;   // at most 24 classes per m2
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) CS(b, lq + 4 * v, l16) = acc[v];
    }
    // stage-B operands: the i block of a wave's outputs is the same for all its
    // blocks (NW is a multiple of nib), so one TB slice serves them all
    // All loads below are unconditional from clamped (in-range) addresses,
    // with out-of-range values zeroed afterwards: loads under a branch would
    // make the first stage-B MFMA wait for every histogram load.
    const int ibw = wave % nib;
    double tb[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) {
        const int kk = 4 * q + lq;
        tb[q] = TB[(int64_t)(16 * ibw + l16) * kb + min(kk, kb - 1)];
    }
#pragma unroll
    for (int q = 0; q < QB; ++q)
        if (4 * q + lq >= kb) tb[q] = 0.0;
    __syncthreads();
    // ---- stage B: blocks (j local, i block), K = 2 n_m2 (padded to kb), computed
    // transposed (rows = columns kt, cols = i) so that a lane's c3 values run
    // along contiguous rows r of the histogram column
    double sp[4] = {0.0, 0.0, 0.0, 0.0}, sq[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < MAXB; ++t) {
        const int ob = wave + NW * t;
        if (ob >= nob) break;
        const int jl = ob / nib, ib = ob % nib;
        const int j = j0 + jl;
        if (j >= NJ) continue;   // wave-uniform
        // histogram values of this block (transposed: lane -> column c0 + lq + 4v,
        // row i = 16 ib + l16), requested before its MFMAs
        double qv[4];
        if (ENERGY) {
            const int iq = min(16 * ib + l16, NI - 1);
#pragma unroll
            for (int v = 0; v < 4; ++v) qv[v] = qbase[(int64_t)max(0, min(lq + 4 * v, nv - 1)) * RQ + iq + NI * j];
        }
        dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < QB; ++q)
            if (4 * q < kb) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(CS(4 * q + lq, jl, l16), tb[q], acc, 0, 0, 0);
        const int i = 16 * ib + l16;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int col = c0 + lq + 4 * v;
            if (lq + 4 * v < nv && i < NI) {
                const int r = i + N * j;
                const double c = acc[v];
                if (OUT == 1) out[(int64_t)col * ldo + r] = c;
                if (ENERGY) {
                    const double pv = c > 1e-30 ? c : 1e-30;
                    sp[v] += pv;
                    sq[v] += qv[v] * tp_fastlog_e(pv, s_lt, lk);
                }
            }
        }
    }
    if (ENERGY) {
        // column c0 + lq + 4v: sum over the 16 lanes of the same lq
#pragma unroll
        for (int v = 0; v < 4; ++v) {
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                sp[v] += __shfl_xor(sp[v], o);
                sq[v] += __shfl_xor(sq[v], o);
            }
            if (l16 == 0) {
                red[wave][0][lq + 4 * v] = sp[v];
                red[wave][1][lq + 4 * v] = sq[v];
            }
        }
        __syncthreads();
        if (tid < 16 && tid < nv) {
            double a = 0.0, bsum = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                a += red[w][0][tid];
                bsum += red[w][1][tid];
            }
            part[((int64_t)jc * 2 + 0) * ldp + c0 + tid] = FOLD ? 4.0 * a : a;   // FOLD: four bins per orbit
            part[((int64_t)jc * 2 + 1) * ldp + c0 + tid] = bsum;
        }
    }
}

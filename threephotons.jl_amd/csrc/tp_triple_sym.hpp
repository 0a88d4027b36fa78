// tp_triple_sym.hpp -- FP64 triple products on sorted l-triples (included by tp_hip.hip).
//
// For even l1 + l2 + l3 the 3j symbol is unchanged by any permutation of its
// columns, so
//   As(k1,k2,k3; l1,l2,l3) = sum 3j(l1 m1 l2 m2 l3 m3) Re(c_k1,l1,m1 c_k2,l2,m2 c_k3,l3,m3)
// (correlations.jl:165-174 with the weights of :116) is the same number for
// every simultaneous permutation of (k1,k2,k3) and (l1,l2,l3).  The kernel
// therefore works on the sorted l-triples s = (L1 >= L2 >= L3) only (125 of
// the 505 at L = 18) and forms, for ALL ordered shell triples,
//   V_s[K1][K2][K3] = Re sum_mu3 T[K1][K2][mu3] c_K3[L3,mu3],
//   T[K1][K2][mu3]  = sum_m1 w(L1 m1 L2 m2 L3 mu3) c_K1[L1,m1] c_K2[L2,m2],  m2 = -m1-mu3,
// then writes As[lt][kt] = V_s[k_sigma^-1(1)..] for each ordered l-triple lt
// = sigma(s) and each sorted k-triple kt.  Full 16 x 16 (K1, K2) tiles: no
// triangular k1 >= k2 >= k3 waste in the MFMA tiles, and T (the stage that
// costs most) is formed once per sorted l-triple instead of once per
// ordering (0.57 GFLOP of FP64 MFMA work per full-size evaluation against
// 1.50 for k_triple_mfma).
//
// Workgroup = (sorted l-triple, K1 half h1, K2 half h2): 8 waves, two
// workgroups per CU.  Per chunk of 8 mu3 groups (one per wave, the groups of
// a chunk of similar size):
//   stage 1: Tr, Ti (16 x 16 each) over the group's E entries, four MFMAs
//            per fetch of four entries (w Re a . Re b, w Re a . Im b,
//            w Im a . -Im b, w Im a . Re b); Tr, Ti -> LDS;
//   stage 2: V(256 pairs x 32 K3) += [Tr | Ti] . [Re c_K3 ; -Im c_K3] over
//            the chunk's 16 slots, in MFMA registers across chunks.
// Then V goes through LDS to the As rows of the l-triple's orderings, each
// ordering through a host table of the tile entries it stores.
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>

constexpr int kSymWaves = 16;                // waves per workgroup: also the mu3 groups of a stage-1 pass
constexpr int kSymThreads = 64 * kSymWaves;
// Batches: 8 waves, so that two workgroups share a CU (measured on the
// 1,280-evaluation sweep: 16.3 ms against 23.0 ms with 16 waves; one
// evaluation is faster with 16 waves, 35 us against 39 us).
constexpr int kSymWavesBatch = 8;
constexpr int kSymThreadsBatch = 64 * kSymWavesBatch;
constexpr bool kSymGauss = true;             // stage 1 with three real products per complex one
// The six permutations sigma of (0, 1, 2): an ordered l-triple (l1, l2, l3) =
// (L_sigma(0), L_sigma(1), L_sigma(2)) of the sorted one.
constexpr int kSymPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// Dynamic LDS of the sorted-triple kernels for order bound Lw = 2L + 1, K
// shells and at most Emax entries per l-triple: FP64 staging, or (s16, the
// mixed mode) FP32 staging with T chunks of NW groups and V staged in FP32.
static inline size_t triple_sym_lds(int K, int Lw, int Emax, int NW = kSymWaves, bool s16 = false)
{
    if (!s16)
        return 16 * sizeof(double2) * 2 * (size_t)Lw + sizeof(double2) * (size_t)K * Lw +
               (sizeof(double) + sizeof(uint32_t)) * (size_t)Emax + sizeof(int) * (2 * (size_t)Lw + 2) +
               sizeof(double) * 2 * NW * 256 + 64;
    const size_t tf = std::max((size_t)NW * 256 * sizeof(uint2), (size_t)256 * 16 * sizeof(float));
    return 16 * sizeof(float2) * 2 * (size_t)Lw + sizeof(float2) * (size_t)K * Lw +
           (sizeof(float) + sizeof(uint16_t)) * (size_t)Emax + sizeof(int) * (2 * (size_t)Lw + 2) + tf + 32;
}

// A work item: everything its operand loads need, in one 32-byte record, so
// that they all issue after one (uniform) load: ts = index of the sorted
// l-triple in the ordered list, its l's and the item's K1 / K2 halves packed
// in tsx (sym_tsx), its entries ew/eidx[e0 .. e0 + ne) and mu3 groups from
// goff (gstart, gorder); e and K: the evaluation of a batch and its shells.
struct SymItem {
    int e, tsx, e0, ne, goff, K, pad0, pad1;
};
__host__ __device__ constexpr int sym_tsx(int ts, int h1, int h2, int L1, int L2, int L3)
{
    return ts | h1 << 11 | h2 << 12 | L1 << 13 | L2 << 18 | L3 << 23;   // ts < 2048, l <= 31
}

// perms[ts * 8] = number n of orderings, perms[ts * 8 + 1 + q]
// = lt | p << 24 with p the index of the permutation sigma of ordering q in
// kSymPerm (k_i = K_sigma(i)); gorder[first group of ts + u] = the mu3 group
// run u-th; symtab[symoff[((p * 2 + h1) * 2 + h2) * 2 + cb] ..
// symoff[.. + 1]) = the V entries of K3 block cb that ordering p stores:
// (K1 - 16 h1 - 1) * 256 + (K2 - 16 h2 - 1) * 16 + (K3 - 16 cb - 1) | kt << 12,
// in kt order.
// The body of one work item wk on coefficients coef of K shells; As entry
// (lt, kt) of the evaluation at As[lt * ldA + as_col(kt, cbase, nprop, p0)],
// entries kt >= klen (shells past K) skipped.
//
// S16 (the mixed-precision mode TP_PREC_FP16X3_T64): stage 1 on FP32 MFMA
// (v_mfma_f32_16x16x4f32, the FP64 form's operand shape at twice the FP64
// rate; c_K1, c_K2 and the weights staged in FP32, T accumulated in FP32,
// rows 4 (lane / 16) + v), stage 2 on FP16 MFMA.
// T (FP64, from stage 1) and c_K3 are scaled by powers of two from the
// l-block norms of the work item's own shells, r_a(K1) = 2^-(1 + ilogb |c_K1,L1|)
// and so on, so that |T| r_a r_b < 1 (the Clebsch-Gordan map is unitary:
// |T(K1,K2,mu3)| <= |c_K1,L1| |c_K2,L2|) and |c_K3| r_c < 1; both are taken
// times 2^13, split into FP16 hi + lo (22 significant bits), and V
// accumulates hi.hi, hi.lo and lo.hi in FP32 on v_mfma_f32_16x16x16f16, one
// K step per 8 mu3 groups (16 slots: Re T and Im T) of a chunk of NW groups
// (one per wave).  V is unscaled exactly (powers
// of two) and staged in FP32 on its way to As.  LDS: c_K3 staged in FP32 and
// split in place ({hi Re | hi -Im << 16, lo Re | lo -Im << 16} per entry);
// T of a chunk as 8-byte {hi Re | hi Im << 16, lo Re | lo Im << 16} per
// (pair, group), the 16-byte pairs of groups (2j, 2j + 1) of pair p at
// position j ^ ((p >> SH) & (NW/2 - 1)) (SH = 1 for 16 groups, 2 for 8), so
// that the stage-1 writes (one group, 64 pairs) and the stage-2 reads (16
// pairs per lane group) both spread over the banks; an entry's local m1 and
// m2 indices in 2 bytes, the item's scales as exponents.  About 40 KB at the
// full size for 8 waves: four workgroups per CU.
template <int NW, bool G3, bool S16 = false, bool A32 = false>
__device__ __forceinline__ void triple_sym_body(double2 *sm, const double2 *__restrict__ coef, int n_c, int K,
                                                const SymItem &it, const double *__restrict__ ew,
                                                const uint32_t *__restrict__ eidx, const int *__restrict__ gstart,
                                                const int *__restrict__ perms, const int *__restrict__ gorder,
                                                const uint32_t *__restrict__ symtab, const int *__restrict__ symoff,
                                                int Lw, double *__restrict__ As, int64_t ldA, int cbase, int nprop,
                                                int p0, uint32_t klen, double *__restrict__ nrm2,
                                                float *__restrict__ As32 = nullptr)
{
    const int ts = it.tsx & 2047, h1 = (it.tsx >> 11) & 1, h2 = (it.tsx >> 12) & 1;
    const int L1 = (it.tsx >> 13) & 31, L2 = (it.tsx >> 18) & 31, L3 = (it.tsx >> 23) & 31;
    const int n1 = 2 * L1 + 1, n2 = 2 * L2 + 1, n3 = 2 * L3 + 1;
    const int goff = it.goff;
    // LDS carve-up (sizes from the order bound Lw so that the layout is fixed;
    // pointers formed on sm itself, no integer round trip, so that their
    // accesses stay LDS accesses instead of flat ones).  FP64: Ab, Bb, Cb
    // double2 and Ws double; S16: their FP32 copies (Cb then split in place).
    constexpr size_t CE = S16 ? sizeof(float2) : sizeof(double2), WE = S16 ? sizeof(float) : sizeof(double);
    char *const pA = reinterpret_cast<char *>(sm);
    char *const pB = pA + 16 * CE * (size_t)Lw;
    char *const pC = pB + 16 * CE * (size_t)Lw;
    char *const pW = pC + CE * (size_t)K * Lw;
    double2 *Ab = reinterpret_cast<double2 *>(pA);         // [n1][16]  c_K1[L1, m1], K1 = 16 h1 + i + 1
    double2 *Bb = reinterpret_cast<double2 *>(pB);         // [n2][16]  c_K2[L2, m2]
    double2 *Cb = reinterpret_cast<double2 *>(pC);         // [n3][K]   c_K3[L3, mu3]
    double *Ws = reinterpret_cast<double *>(pW);           // [E] weights
    float2 *Abf = reinterpret_cast<float2 *>(pA), *Bbf = reinterpret_cast<float2 *>(pB);   // S16
    float2 *Cf = reinterpret_cast<float2 *>(pC);
    uint2 *Cu = reinterpret_cast<uint2 *>(pC);             // S16: c_K3 split, [n3][K]
    float *Wsf = reinterpret_cast<float *>(pW);
    const int e0 = it.e0, ne = it.ne;
    uint32_t *Is = reinterpret_cast<uint32_t *>(pW + WE * (size_t)ne);   // [E] local (m1 index | m2 index << 16)
    uint16_t *Is16 = reinterpret_cast<uint16_t *>(pW + WE * (size_t)ne); // S16: [E] m1 index | m2 index << 8
    constexpr size_t IE = S16 ? sizeof(uint16_t) : sizeof(uint32_t);
    int *Gs = reinterpret_cast<int *>(pW + (WE + IE) * (size_t)ne + (S16 ? 2 * (ne & 1) : 0));   // [n3 + 1] group starts
    int *Od = Gs + (n3 + 1);                                             // [n3] run order of the groups
    // FP64: [2 NW][256] T slots, then the V staging; S16: T chunk, then the V staging in FP32
    const size_t tf_off = (reinterpret_cast<const char *>(Od + n3) - pA + 15) & ~(size_t)15;
    double *Tf = reinterpret_cast<double *>(pA + tf_off);
    uint2 *T2 = reinterpret_cast<uint2 *>(pA + tf_off);
    float *Tff = reinterpret_cast<float *>(pA + tf_off);
    constexpr int NT = 64 * NW, NR = 16 / NW;   // threads; stage-2 row blocks per wave

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int co1 = L1 * (L1 - 1) / 2, co2 = L2 * (L2 - 1) / 2, co3 = L3 * (L3 - 1) / 2;   // compact offsets (even l)
    const int K1b = 16 * h1, K2b = 16 * h2;
    for (int x = tid; x < 16 * n1; x += NT) {
        const int m = x >> 4, i = x & 15, k = K1b + i;
        const double2 c = k < K ? coef[(size_t)k * n_c + co1 + m] : make_double2(0.0, 0.0);
        if constexpr (S16) Abf[x] = make_float2((float)c.x, (float)c.y);
        else Ab[x] = c;
    }
    for (int x = tid; x < 16 * n2; x += NT) {
        const int m = x >> 4, j = x & 15, k = K2b + j;
        const double2 c = k < K ? coef[(size_t)k * n_c + co2 + m] : make_double2(0.0, 0.0);
        if constexpr (S16) Bbf[x] = make_float2((float)c.x, (float)c.y);
        else Bb[x] = c;
    }
    for (int x = tid; x < K * n3; x += NT) {
        const int m = x / K, k = x - m * K;
        const double2 c = coef[(size_t)k * n_c + co3 + m];
        if constexpr (S16) Cf[x] = make_float2((float)c.x, (float)c.y);
        else Cb[x] = c;
    }
    for (int x = tid; x < ne; x += NT) {
        if constexpr (S16) Wsf[x] = (float)ew[e0 + x];
        else Ws[x] = ew[e0 + x];
        const uint32_t id = eidx[e0 + x];
        if constexpr (S16)
            Is16[x] = (uint16_t)(((int)(id & 0xffffu) - co1) | (((int)(id >> 16) - co2) << 8));   // local indices < 2L + 1 <= 45
        else
            Is[x] = (uint32_t)((int)(id & 0xffffu) - co1) | ((uint32_t)((int)(id >> 16) - co2) << 16);
    }
    for (int x = tid; x <= n3; x += NT) Gs[x] = gstart[goff + x] - e0;
    for (int x = tid; x < n3; x += NT) Od[x] = gorder[goff + x];
    __syncthreads();
    // the work item (l, l, l) with h1 = h2 = 0 holds the l-blocks of every
    // shell: their squared norms, nrm2[k][l/2] (column scales of the FP16
    // contraction), from the FP64 coefficients
    if (nrm2 && h1 == 0 && h2 == 0 && L1 == L3 && L2 == L3)
        for (int k = tid; k < K; k += NT) {
            double s2 = 0.0;
            for (int m = 0; m < n3; ++m) {
                const double2 c = S16 ? coef[(size_t)k * n_c + co3 + m] : Cb[m * K + k];
                s2 += c.x * c.x + c.y * c.y;
            }
            nrm2[k * kNrmLd + L3 / 2] = s2;
        }

    const int ncb = K > 16 ? 2 : 1;                 // K3 column blocks
    // S16: the work item's scales r_a (K1 = K1b + i), r_b (K2 = K2b + j), r_c (K3) = 2^-s_ex
    __shared__ int s_ex[S16 ? 64 : 1];
    if constexpr (S16) {
        // squared norms of the 32 + K shell blocks (K1 block, K2 block, K3),
        // TPC consecutive lanes per block, each summing every TPC-th m
        constexpr int TPC = NT / 64;
        const int col = tid / TPC, sub = tid % TPC;
        const int n = col < 16 ? n1 : col < 32 ? n2 : n3;
        double s2 = 0.0;
        if (col < 32 + K)
            for (int m = sub; m < n; m += TPC) {
                const float2 c = col < 16 ? Abf[m * 16 + col] : col < 32 ? Bbf[m * 16 + col - 16] : Cf[m * K + col - 32];
                s2 += (double)c.x * c.x + (double)c.y * c.y;
            }
#pragma unroll
        for (int o = 1; o < TPC; o <<= 1) s2 += __shfl_xor(s2, o);
        if (sub == 0 && col < 32 + K) s_ex[col] = s2 > 0.0 ? 1 + ilogb(sqrt(s2)) : 0;
        __syncthreads();
        // c_K3 (r_c 2^13) split in place, entry by entry (the same 8 bytes)
        for (int x = tid; x < n3 * K; x += NT) {
            const float2 c = Cf[x];
            const double sc = ldexp(1.0, 13 - s_ex[32 + x % K]);
            uint16_t hr, lr, hi, li;
            f16_split((double)c.x * sc, hr, lr);
            f16_split(-(double)c.y * sc, hi, li);
            Cu[x] = make_uint2(hr | (uint32_t)hi << 16, lr | (uint32_t)li << 16);
        }
        __syncthreads();
    }
    dbl4 acc[NR][2];                                // [row block rb = wave + NW r][K3 block]
    f32x4 a16[NR][2];                               // the same in FP32 (S16)
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            acc[r][c] = dbl4{0.0, 0.0, 0.0, 0.0};
            a16[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
        }

    // S16: the T chunk's swizzle, pair p's 16-byte pair of groups j at j ^ ((p >> SH) & FM)
    constexpr int SH = NW == 16 ? 1 : 2, FM = NW / 2 - 1;
    static_assert(!S16 || NW == 16 || NW == 8, "the FP16 stage 2 takes 16 or 8 groups per chunk");
    const int sw = (l16 >> SH) & FM;   // (p >> SH) & FM for the pairs p = 16 x + l16 a lane touches
    for (int c0 = 0; c0 < n3; c0 += NW) {
        const int nu = min(NW, n3 - c0);
        const int gi = wave;   // this wave's group within the chunk
        // ---- stage 1: this wave's group, k-step = 4 entries; with G3 the
        // complex product as P1 = (w Re a) Re b, P2 = (w Im a) Im b,
        // P3 = (w (Re a + Im a)) (Re b + Im b): Tr = P1 - P2, Ti = P3 - P1 - P2;
        // otherwise Tr += (w Re a) Re b - (w Im a) Im b, Ti += (w Re a) Im b + (w Im a) Re b
        if constexpr (S16) {
            // FP32: rows K1 = 4 lq + v of the accumulators
            f32x4 q1 = f32x4{0.f, 0.f, 0.f, 0.f}, q2 = q1, q3 = q1;
            if (gi < nu) {
                const int u = Od[c0 + gi];
                const int s = Gs[u], E = Gs[u + 1] - s;
                const int nks = (E + 3) >> 2;
#pragma unroll 2
                for (int ks = 0; ks < nks; ++ks) {
                    const int e = 4 * ks + lq;
                    float ar = 0.f, ai = 0.f, br = 0.f, bi = 0.f;
                    if (e < E) {
                        const float w = Wsf[s + e];
                        const uint32_t id = Is16[s + e];
                        const float2 ca = Abf[(int)(id & 0xffu) * 16 + l16];
                        const float2 cbv = Bbf[(int)(id >> 8) * 16 + l16];
                        ar = w * ca.x;
                        ai = w * ca.y;
                        br = cbv.x;
                        bi = cbv.y;
                    }
                    q1 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, br, q1, 0, 0, 0);
                    q2 = __builtin_amdgcn_mfma_f32_16x16x4f32(ai, bi, q2, 0, 0, 0);
                    q3 = __builtin_amdgcn_mfma_f32_16x16x4f32(ar + ai, br + bi, q3, 0, 0, 0);
                }
            }
            // pair p = 16 (4 lq + v) + l16
            const int exb = 13 - s_ex[16 + l16];
            const int gpos = 2 * ((gi >> 1) ^ sw) + (gi & 1);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int p = 16 * (4 * lq + v) + l16;
                const float sc = ldexpf(1.f, exb - s_ex[4 * lq + v]);
                const float tr = (q1[v] - q2[v]) * sc, ti = (q3[v] - q1[v] - q2[v]) * sc;
                const _Float16 hr = (_Float16)tr, hi = (_Float16)ti;
                const _Float16 lr = (_Float16)(tr - (float)hr), li = (_Float16)(ti - (float)hi);
                T2[p * NW + gpos] = make_uint2((uint32_t)__builtin_bit_cast(uint16_t, hr) |
                                                   (uint32_t)__builtin_bit_cast(uint16_t, hi) << 16,
                                               (uint32_t)__builtin_bit_cast(uint16_t, lr) |
                                                   (uint32_t)__builtin_bit_cast(uint16_t, li) << 16);
            }
        } else {
            dbl4 p1 = dbl4{0.0, 0.0, 0.0, 0.0}, p2 = dbl4{0.0, 0.0, 0.0, 0.0}, p3 = dbl4{0.0, 0.0, 0.0, 0.0};
            if (gi < nu) {
                const int u = Od[c0 + gi];
                const int s = Gs[u], E = Gs[u + 1] - s;
                const int nks = (E + 3) >> 2;
#pragma unroll 2
                for (int ks = 0; ks < nks; ++ks) {
                    const int e = 4 * ks + lq;
                    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
                    if (e < E) {
                        const double w = Ws[s + e];
                        const uint32_t id = Is[s + e];
                        const double2 ca = Ab[(int)(id & 0xffffu) * 16 + l16];
                        const double2 cbv = Bb[(int)(id >> 16) * 16 + l16];
                        ar = w * ca.x;
                        ai = w * ca.y;
                        br = cbv.x;
                        bi = cbv.y;
                    }
                    if (G3) {
                        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, p1, 0, 0, 0);
                        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, bi, p2, 0, 0, 0);
                        p3 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar + ai, br + bi, p3, 0, 0, 0);
                    } else {
                        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, p1, 0, 0, 0);
                        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, p2, 0, 0, 0);
                        p1 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, -bi, p1, 0, 0, 0);
                        p2 = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, p2, 0, 0, 0);
                    }
                }
            }
            // rows K1 = lq + 4 v, columns K2 = l16: pair p = 16 row + column
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int p = 16 * (lq + 4 * v) + l16;
                Tf[(2 * gi) * 256 + p] = G3 ? p1[v] - p2[v] : p1[v];
                Tf[(2 * gi + 1) * 256 + p] = G3 ? p3[v] - p1[v] - p2[v] : p2[v];
            }
        }
        __syncthreads();
        if constexpr (S16) {
            // ---- stage 2 on FP16 MFMA (v_mfma_f32_16x16x16f16): each 8 groups of
            // the chunk (16 slots) are one K step, lane (l16, lq) holding slots
            // 4 lq .. 4 lq + 3 = (Re, Im) of groups 2 lq, 2 lq + 1 of the 8; B =
            // [Re c_K3 ; -Im c_K3] of those groups, A = T of pair 16 rb + l16.
            // With 16 waves the two halves of a chunk run in this order, so every
            // accumulator sees the same MFMA sequence as with 8 waves (batches
            // give the energies of single evaluations bit for bit).
#pragma unroll
            for (int cb = 0; cb < 2; ++cb) {
                if (cb >= ncb) break;
                const int k3 = 16 * cb + l16;
#pragma unroll
                for (int hh = 0; hh < NW / 8; ++hh) {
                    u32x2 bh, bl;
#pragma unroll
                    for (int q = 0; q < 2; ++q) {
                        const int g = 8 * hh + 2 * lq + q;
                        const uint2 u = g < nu && k3 < K ? Cu[Od[c0 + g] * K + k3] : make_uint2(0u, 0u);
                        bh[q] = u.x;
                        bl[q] = u.y;
                    }
                    const h16x4 hbh = __builtin_bit_cast(h16x4, bh), hbl = __builtin_bit_cast(h16x4, bl);
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
                        const uint2 *trow = T2 + (16 * (wave + NW * r) + l16) * NW;
                        const uint4 q1 = *reinterpret_cast<const uint4 *>(trow + 2 * ((4 * hh + lq) ^ sw));
                        const h16x4 ah = __builtin_bit_cast(h16x4, u32x2{q1.x, q1.z});
                        const h16x4 al = __builtin_bit_cast(h16x4, u32x2{q1.y, q1.w});
                        a16[r][cb] = __builtin_amdgcn_mfma_f32_16x16x16f16(al, hbh, a16[r][cb], 0, 0, 0);
                        a16[r][cb] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, hbl, a16[r][cb], 0, 0, 0);
                        a16[r][cb] = __builtin_amdgcn_mfma_f32_16x16x16f16(ah, hbh, a16[r][cb], 0, 0, 0);
                    }
                }
            }
            __syncthreads();
            continue;
        }
        // ---- stage 2: slots 2u' (Re T) and 2u'+1 (Im T) of this chunk
        const int nsl = 2 * nu, nks2 = (nsl + 3) >> 2;
        for (int ks = 0; ks < nks2; ++ks) {
            const int slot = 4 * ks + lq;
            const bool ok = slot < nsl;
            double bv[2] = {0.0, 0.0};
            if (ok) {
                const int m = Od[c0 + (slot >> 1)];
#pragma unroll
                for (int cb = 0; cb < 2; ++cb) {
                    const int k3 = 16 * cb + l16;
                    if (cb < ncb && k3 < K) {
                        const double2 c = Cb[m * K + k3];
                        bv[cb] = (slot & 1) ? -c.y : c.x;
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double av = ok ? Tf[slot * 256 + 16 * (wave + NW * r) + l16] : 0.0;
                acc[r][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[0], acc[r][0], 0, 0, 0);
                if (ncb > 1) acc[r][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[1], acc[r][1], 0, 0, 0);
            }
        }
        __syncthreads();
    }

    // ---- V to the As rows of every ordering of the l-triple (tables of the
    // entries each ordering stores, in k-triple order).  With L1 = L2, T and
    // V are symmetric in (K1, K2) (the 3j symbol is unchanged by exchanging
    // its first two columns, l1 + l2 + l3 even), so the work lists hold no
    // item with h1 < h2: item (h1, h2), h1 > h2, also stores the entries of
    // tile (h2, h1), V(K1, K2, K3) = V(K2, K1, K3) read at the exchanged
    // local index.
    const int np = perms[ts * 8];
    const bool mirror = L1 == L2 && h1 > h2;
    auto vstore = [&](int64_t row, uint32_t kt, uint32_t idx) {
        const int64_t col = as_col(kt, cbase, nprop, p0);
        if constexpr (!S16) As[row + col] = Tf[idx];
        else if constexpr (A32) As32[row + col] = Tff[idx];
        else As[row + col] = (double)Tff[idx];
    };
    for (int cb = 0; cb < ncb; ++cb) {
        // V tile of K3 block cb: Vs[pair][k3 - 16 cb] (pairs 16 rb + lq + 4 v, k3 = l16;
        // S16: pairs 16 rb + 4 lq + v, unscaled by 2^-26 / (r_a r_b r_c), in FP32)
#pragma unroll
        for (int r = 0; r < NR; ++r)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if constexpr (S16) {
                    const int p = 16 * (wave + NW * r) + 4 * lq + v, k3 = 16 * cb + l16;
                    const double un =
                        k3 < K ? ldexp(1.0, s_ex[p >> 4] + s_ex[16 + (p & 15)] + s_ex[32 + k3] - 26) : 0.0;
                    Tff[p * 16 + l16] = (float)((double)a16[r][cb][v] * un);
                } else {
                    Tf[(16 * (wave + NW * r) + lq + 4 * v) * 16 + l16] = acc[r][cb][v];
                }
            }
        __syncthreads();
        for (int q = 0; q < np; ++q) {
            const int code = perms[ts * 8 + 1 + q];
            const int64_t row = (int64_t)(code & 0xffffff) * ldA;
            const int t = (((code >> 24) * 2 + h1) * 2 + h2) * 2 + cb;
            const int o = symoff[t], n = symoff[t + 1] - o;
            for (int x = tid; x < n; x += NT) {
                const uint32_t ent = symtab[o + x], kt = ent >> 12;
                if (kt < klen) vstore(row, kt, ent & 4095u);
            }
            if (mirror) {
                const int tm = (((code >> 24) * 2 + h2) * 2 + h1) * 2 + cb;
                const int om = symoff[tm], nm = symoff[tm + 1] - om;
                for (int x = tid; x < nm; x += NT) {
                    const uint32_t ent = symtab[om + x], kt = ent >> 12, e = ent & 4095u;
                    const uint32_t ex = ((e >> 4) & 15u) * 256u + (e >> 8) * 16u + (e & 15u);
                    if (kt < klen) vstore(row, kt, ex);
                }
            }
        }
        __syncthreads();
    }
}

template <int NW, bool G3, bool S16 = false>
__global__ __launch_bounds__(64 * NW) void k_triple_sym_t(const double2 *__restrict__ cb0,
                                                           const double2 *__restrict__ cb1,
                                                           const int *__restrict__ sel, int flip, int n_c, int K,
                                                           const double *__restrict__ ew,
                                                           const uint32_t *__restrict__ eidx,
                                                           const int *__restrict__ gstart,
                                                           const SymItem *__restrict__ items,
                                                           const int *__restrict__ perms,
                                                           const int *__restrict__ gorder,
                                                           const uint32_t *__restrict__ symtab,
                                                           const int *__restrict__ symoff, int Lw,
                                                           double *__restrict__ As, int64_t ldA,
                                                           double *__restrict__ nrm2)
{
    extern __shared__ double2 sm[];
    const SymItem it = items[blockIdx.x];
    triple_sym_body<NW, G3, S16>(sm, pick(cb0, cb1, sel, flip), n_c, K, it, ew, eidx, gstart, perms, gorder, symtab,
                                 symoff, Lw, As, ldA, 0, 1, 0, 0xffffffffu, nrm2);
}

// The same on a batch of evaluations (struct EvalDesc): item.e is the
// evaluation (coefficients at coef + e cstride), item.K its shells; As in FP32
// (its only reader, k_hgemm_g, rounds As x (a power of two) to FP32 before
// the FP16 split, which gives the same bits from the FP32 value); the
// tables are those of the batch's basis (K_b shells): the entries of an
// evaluation of K <= K_b shells are the ones with kt < K (K + 1) (K + 2) / 6,
// since a sorted k-triple's index does not depend on the shell count.
template <int NW, bool G3, bool S16, typename Desc>
__global__ __launch_bounds__(64 * NW, S16 ? 4 : 2) void k_triple_sym_batch_t(const Desc *__restrict__ ev,
                                                                 const SymItem *__restrict__ items,
                                                                 const double2 *__restrict__ coef, int64_t cstride,
                                                                 int n_c, const double *__restrict__ ew,
                                                                 const uint32_t *__restrict__ eidx,
                                                                 const int *__restrict__ gstart,
                                                                 const int *__restrict__ perms,
                                                                 const int *__restrict__ gorder,
                                                                 const uint32_t *__restrict__ symtab,
                                                                 const int *__restrict__ symoff, int Lw,
                                                                 float *__restrict__ As32, int64_t ldA)
{
    extern __shared__ double2 sm[];
    const SymItem it = items[blockIdx.x];
    if (it.e < 0) return;       // padding of an XCD queue (whole workgroup, before any barrier)
    const Desc &d = ev[it.e];   // column placement, needed only at the write-out
    const int K = it.K;
    triple_sym_body<NW, G3, S16, true>(sm, coef + it.e * cstride, n_c, K, it, ew, eidx, gstart, perms, gorder,
                                       symtab, symoff, Lw, nullptr, ldA, d.cbase, d.nprop, d.p,
                                       (uint32_t)(K * (K + 1) * (K + 2) / 6), nullptr, As32);
}

// The configuration the library launches.
#define k_triple_sym (k_triple_sym_t<kSymWaves, kSymGauss>)
// TP_PREC_FP16X3_T64 takes stage 2 on FP16 MFMA, in batches (14.9 ms against
// 15.6 ms per 1,280-evaluation sweep with the FP64 stage 2) and, so that a
// batch gives what single evaluations give, in single evaluations.
#define k_triple_sym16 (k_triple_sym_t<kSymWaves, kSymGauss, true>)
#define k_triple_sym_batch (k_triple_sym_batch_t<kSymWavesBatch, kSymGauss, true, EvalDesc>)

// tp_hgemm.hpp -- G = H . As on FP16 hi/lo MFMA operands (included by tp_hip.hip
// after tp_kernels.hpp).
//
// The contraction c3 = Bm . (H . As) of the reference's B . PA
// (correlations.jl:179-193) in mixed precision: the class weights H
// [class][l-triple] (FP64 on the host, scaled by one power of two into
// [2^13, 2^14) and split into FP16 hi + lo) times the FP64 triple products
// As [l-triple][k-triple], which the kernel scales per column by a power of
// two and splits into hi + lo while staging them into LDS (no 16-bit copy of
// As in HBM).  Products hi.hi and hi.lo + lo.hi go to two FP32 accumulators
// (v_mfma_f32_16x16x32_f16), summed in FP64 and unscaled in the epilogue: G
// is FP64 [class][column], the input of the separated-grid contraction
// k_c3sep, which stays in FP64.  2 n_m n_lt flops per column (0.137 MFLOP at
// L = 18) instead of the dense F . As's 2 R n_lt (2.07 MFLOP).
//
// Workgroup: BN columns x NRB row blocks of 16 classes (blockIdx.y selects
// the row blocks [NRB y, NRB y + NRB); H has zero rows up to kHgRows), 4
// waves; wave w owns column block w % (BN/16) and every (4 / (BN/16))-th row
// block from w / (BN/16), the same number for every wave (NRB is a multiple
// of 4 / (BN/16)), so the MFMA loop has no branches.  K (l-triples) in
// slices of KS staged through LDS, the next slice prefetched into registers
// (KS = 128 for one evaluation: few slices, so little serial load latency).
// Column scales: from the shell norms (SINGLE: one evaluation,
// nrm2[k][l/2] = |c_k,l|^2 written by k_triple_sym) or from sc[] (batches:
// k_colscale_batch).  Batches (A32) keep As in FP32: f16_split rounds its
// argument to FP32 first, and a power-of-two scale commutes with that
// rounding, so the split of (double) As32 x sc equals that of As x sc.
#pragma once

#include "tp_kernels.hpp"

constexpr int kHgRows = 256;   // rows of the split H (zero past the classes): 16 row blocks
constexpr int kHgMaxRb = 12;   // row blocks of classes supported: n_m <= 192 (L <= 20)

// Power-of-two scale of k-triple column kt from the shell norms nk (K <= 64):
// |As(kt, lt)| <= |c_k1,l1| |c_k2,l2| |c_k3,l3| <= nk1 nk2 nk3 (the 3j
// coupling of three unit vectors has modulus <= 1), so |As s| < 2^14.
__device__ __forceinline__ double hg_kt_scale(int kt, const double *nk)
{
    int k1 = 1;
    while ((k1 + 1) * (k1 + 2) * k1 / 6 <= kt) ++k1;
    const int r1 = kt - (k1 - 1) * k1 * (k1 + 1) / 6;
    int k2 = 1;
    while ((k2 + 1) * k2 / 2 <= r1) ++k2;
    const int k3 = r1 - (k2 - 1) * k2 / 2 + 1;
    const double bnd = nk[k1 - 1] * nk[k2 - 1] * nk[k3 - 1];
    return bnd > 0.0 ? ldexp(1.0, 13 - ilogb(bnd)) : 1.0;
}

template <int BN, int NRB, bool SINGLE, int KS = 32, bool A32 = false>
__global__ __launch_bounds__(256) void k_hgemm_g(const uint16_t *__restrict__ Hhi, const uint16_t *__restrict__ Hlo,
                                                 int ldh, double h_inv, const double *__restrict__ As, int64_t ldA,
                                                 int nlt, int kpad, const double *__restrict__ sc,
                                                 const double *__restrict__ nrm2, int L, int K, int grows,
                                                 double *__restrict__ G, int64_t ldg,
                                                 const float *__restrict__ As32 = nullptr)
{
    constexpr int CBN = BN / 16, WPC = 4 / CBN, NI = NRB / WPC;
    constexpr int NROW = NRB * 16;
    constexpr int ROW = KS + 8;                           // halves per LDS row: 16-byte reads without bank conflicts
    constexpr int NCHA = NROW * (KS / 8) * 2;             // 16-byte A chunks per slice (both planes)
    constexpr int QA = (NCHA + 255) / 256;
    constexpr int NBT = BN * (KS / 8);                    // threads staging B (8 rows each)
    static_assert(BN == 16 || BN == 32 || BN == 64, "BN");
    static_assert(KS % 32 == 0 && NBT <= 256, "slice");
    static_assert(NRB % WPC == 0, "every wave gets the same number of row blocks");
    __shared__ uint16_t sA[2][NROW][ROW];
    __shared__ uint16_t sB[2][BN][ROW];
    __shared__ double s_sc[BN], s_isc[BN];
    __shared__ double s_nk[SINGLE ? 64 : 1];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int64_t c0 = (int64_t)blockIdx.x * BN;
    const int row0 = (int)blockIdx.y * NROW;
    const int cbw = wave % CBN, roff = wave / CBN;

    // ---- slice loads (registers), then LDS
    u16x8 pa[QA];
    double pb[8];
    float pbf[8];   // A32: the prefetched FP32 values, converted where staged (not waited on at the load)
    auto loadA = [&](int k0) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int i = tid + 256 * q;
            if (NCHA % 256 == 0 || i < NCHA) {
                const int p = i / (NROW * (KS / 8)), rem = i - p * NROW * (KS / 8);
                const int row = rem / (KS / 8), kc = (rem % (KS / 8)) * 8;
                const uint16_t *base = p ? Hlo : Hhi;
                pa[q] = *reinterpret_cast<const u16x8 *>(base + (int64_t)(row0 + row) * ldh + k0 + kc);
            }
        }
    };
    const int bc = tid % BN, bg = tid / BN;   // B staging: column bc, rows 8 bg .. 8 bg + 7 of the slice
    auto loadB = [&](int k0) {
        if (tid < NBT) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int lt = k0 + 8 * bg + i;
                if constexpr (A32) pbf[i] = lt < nlt ? As32[(int64_t)lt * ldA + c0 + bc] : 0.f;
                else pb[i] = lt < nlt ? As[(int64_t)lt * ldA + c0 + bc] : 0.0;
            }
        }
    };
    loadA(0);
    loadB(0);

    // ---- column scales
    if (SINGLE) {
        if (tid < K) {
            double m = 0.0;
            for (int l = 0; l <= L / 2; ++l) m = fmax(m, nrm2[tid * kNrmLd + l]);
            s_nk[tid] = sqrt(m);
        }
        __syncthreads();
        const int klength = K * (K + 1) * (K + 2) / 6;
        if (tid < BN) {
            const int kt = (int)c0 + tid;
            const double s = kt < klength ? hg_kt_scale(kt, s_nk) : 1.0;
            s_sc[tid] = s;
            s_isc[tid] = 1.0 / s;
        }
    } else if (tid < BN) {
        double s = sc[c0 + tid];
        if (!(s > 0.0 && s < 1e300)) s = 1.0;   // columns no evaluation owns (zero As)
        s_sc[tid] = s;
        s_isc[tid] = 1.0 / s;
    }
    __syncthreads();

    f32x4 acc[NI], cor[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
        cor[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    const double bsc = s_sc[bc];
    for (int k0 = 0; k0 < kpad; k0 += KS) {
        __syncthreads();   // the previous slice's LDS reads are done
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int i = tid + 256 * q;
            if (NCHA % 256 == 0 || i < NCHA) {
                const int p = i / (NROW * (KS / 8)), rem = i - p * NROW * (KS / 8);
                const int row = rem / (KS / 8), kc = (rem % (KS / 8)) * 8;
                *reinterpret_cast<u16x8 *>(&sA[p][row][kc]) = pa[q];
            }
        }
        if (tid < NBT) {
            u16x8 h, l;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                uint16_t hb, lb;
                f16_split((A32 ? (double)pbf[i] : pb[i]) * bsc, hb, lb);
                h[i] = hb;
                l[i] = lb;
            }
            *reinterpret_cast<u16x8 *>(&sB[0][bc][8 * bg]) = h;
            *reinterpret_cast<u16x8 *>(&sB[1][bc][8 * bg]) = l;
        }
        __syncthreads();
        if (k0 + KS < kpad) {
            loadA(k0 + KS);
            loadB(k0 + KS);
        }
#pragma unroll
        for (int kk = 0; kk < KS; kk += 32) {
            const int kq = kk + 8 * lq;
            const h16x8 bh = __builtin_bit_cast(h16x8, *reinterpret_cast<const u16x8 *>(&sB[0][cbw * 16 + l16][kq]));
            const h16x8 bl = __builtin_bit_cast(h16x8, *reinterpret_cast<const u16x8 *>(&sB[1][cbw * 16 + l16][kq]));
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                const int rl = roff + i * WPC;   // row block within the workgroup's rows
                const h16x8 ah = __builtin_bit_cast(h16x8, *reinterpret_cast<const u16x8 *>(&sA[0][rl * 16 + l16][kq]));
                const h16x8 al = __builtin_bit_cast(h16x8, *reinterpret_cast<const u16x8 *>(&sA[1][rl * 16 + l16][kq]));
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[i], 0, 0, 0);
                cor[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, cor[i], 0, 0, 0);
                cor[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, cor[i], 0, 0, 0);
            }
        }
    }
    // ---- epilogue: rows 4 lq + v of row block rl, column l16 of block cbw;
    // rows at or past grows are not written (H is zero there)
    const int cl = cbw * 16 + l16;
    const double osc = h_inv * s_isc[cl];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int rl = roff + i * WPC;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int cls = row0 + rl * 16 + 4 * lq + v;
            if (cls < grows) G[(int64_t)cls * ldg + c0 + cl] = ((double)acc[i][v] + (double)cor[i][v]) * osc;
        }
    }
}

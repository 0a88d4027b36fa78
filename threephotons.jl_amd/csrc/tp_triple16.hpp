// tp_triple16.hpp -- triple products of the mixed-precision mode on FP16 MFMA
// with hi/lo compensation (included by tp_hip.hip after tp_kernels.hpp).
//
// Same coupling form and work decomposition as k_triple_mfma (workgroup =
// (l-triple, half of the k1 range), 8 waves), with every product split:
//   x = hi + lo, hi = fp16(x), lo = fp16(x - hi)   (22 significant bits)
//   x y ~ hi_x hi_y + hi_x lo_y + lo_x hi_y         (one FP32 accumulator: its
//                                                    rounding, 2^-24 of the sum,
//                                                    is below the 2^-22 of the split)
// on v_mfma_f32_16x16x32_f16 (K-depth 32 per instruction instead of 4).
// Range: shell k is scaled by the power of two s_k with n_k s_k in
// [2^12, 2^13), n_k = max_l |c_k,l|, so every scaled entry is <= 2^13.  By the
// unitarity of the Clebsch-Gordan map, |T_k1k2(mu3)| <= |c_k1,l1| |c_k2,l2|
// <= 2^26 (scaled), which stage 2 takes as T 2^-13 <= 2^13, and
// |As| <= 2^26 (scaled); the result is unscaled in FP64:
//   As = acc 2^13 / (s_k1 s_k2 s_k3).
#pragma once

#include "tp_kernels.hpp"

constexpr int kT16Row = 36;   // Tf row stride in floats: 32 slots (16 mu3 x re/im) + 4 pad

__device__ __forceinline__ void split_h(float x, int j, h16x8 &hi, h16x8 &lo)
{
    const _Float16 h = (_Float16)x;
    hi[j] = h;
    lo[j] = (_Float16)(x - (float)h);
}

// One workgroup per shell k: |c_k,l|^2 per block l in parallel, then the
// power of two s_k (out[k]) and the norm n_k = max_l |c_k,l| (out[K + k],
// read by k_split_as for its column scales).
__global__ __launch_bounds__(256) void k_shell_scales(const double2 *__restrict__ cb0, const double2 *__restrict__ cb1,
                                                      const int *__restrict__ sel, int flip, int n_c, int L,
                                                      double *__restrict__ out)
{
    __shared__ double s_e[1024];
    __shared__ double s_l[32];
    const double2 *c = pick(cb0, cb1, sel, flip) + (size_t)blockIdx.x * n_c;
    const int tid = threadIdx.x, nl = L / 2 + 1;
    for (int i = tid; i < n_c; i += 256) s_e[i] = c[i].x * c[i].x + c[i].y * c[i].y;
    __syncthreads();
    if (tid < nl) {
        const int l = 2 * tid;
        double s = 0.0;
        for (int m = 0; m <= 2 * l; ++m) s += s_e[l * (l - 1) / 2 + m];
        s_l[tid] = s;
    }
    __syncthreads();
    if (tid == 0) {
        double m = 0.0;
        for (int l = 0; l < nl; ++l) m = fmax(m, s_l[l]);
        const double nk = sqrt(m);
        out[blockIdx.x] = nk > 0.0 ? ldexp(1.0, 12 - ilogb(nk)) : 1.0;
        out[gridDim.x + blockIdx.x] = nk;
    }
}

// Bytes of dynamic LDS of triple16_body.
__host__ __device__ inline size_t triple16_lds(int K, int Lw, int Emax, int Prows)
{
    auto r16 = [](size_t x) { return (x + 15) / 16 * 16; };
    size_t o = 64 * sizeof(double);                  // shell scales
    o += r16(3 * (size_t)K * Lw * sizeof(float2));   // c1b, c2b, c3b
    o += (size_t)Prows * kT16Row * sizeof(float);    // Tf
    o += (size_t)Emax * (sizeof(float) + sizeof(uint32_t));
    o += ((size_t)Lw + 1 + Prows) * sizeof(int);
    return r16(o);
}

__device__ __forceinline__ uint32_t split_pack(double x)   // fp16 hi | fp16 lo << 16
{
    const float f = (float)x;
    const _Float16 h = (_Float16)f;
    const _Float16 l = (_Float16)(f - (float)h);
    return (uint32_t)__builtin_bit_cast(uint16_t, h) | ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16);
}

// Entries: per (l-triple, mu3) group padded to a multiple of 8 with zero
// weights (ew8 / eidx8 / gstart8), so a lane's 8 consecutive entries are two
// 16-byte reads and lie in one half (Re / Im part) of the K dimension.
__device__ __forceinline__ void triple16_body(char *smem, const double2 *__restrict__ coef,
                                              const double *__restrict__ shsc, int n_c, int K, int lt, int h,
                                              const float *__restrict__ ew8, const uint32_t *__restrict__ eidx8,
                                              const int *__restrict__ gstart8, const int4 *__restrict__ ltinfo, int Lw,
                                              int Emax, int Prows, double *__restrict__ As, int64_t ldA, int cbase,
                                              int nprop, int p0)
{
    const int4 li = ltinfo[lt];
    const int l3 = li.z, l1 = li.w & 0xff, l2 = (li.w >> 8) & 0xff;
    const int n1 = 2 * l1 + 1, n2 = 2 * l2 + 1, n3 = 2 * l3 + 1;
    const int goff = li.y;
    // LDS carve-up (16-byte aligned pieces where vector reads happen)
    double *ssc = reinterpret_cast<double *>(smem);                          // [64]
    float2 *c1b = reinterpret_cast<float2 *>(smem + 64 * sizeof(double));   // [K][n1] scaled FP32
    uint2 *c2h = reinterpret_cast<uint2 *>(c1b + (size_t)K * Lw);            // [K][n2] {re hi|lo, im hi|lo}
    uint2 *c3h = c2h + (size_t)K * Lw;                                       // [K][n3]
    float *Tf = reinterpret_cast<float *>(smem + 64 * sizeof(double) +
                                          (3 * (size_t)K * Lw * sizeof(float2) + 15) / 16 * 16);   // [Prows][36]
    float *sw = Tf + (size_t)Prows * kT16Row;                                // [Emax] (16-byte aligned)
    uint32_t *si = reinterpret_cast<uint32_t *>(sw + Emax);                  // [Emax]
    int *sg = reinterpret_cast<int *>(si + Emax);                            // [n3 + 1]
    int *pk = sg + (Lw + 1);                                                 // [Prows]

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l16 = lane & 15, lq = lane >> 4;
    const int co1 = l1 * (l1 - 1) / 2, co2 = l2 * (l2 - 1) / 2, co3 = l3 * (l3 - 1) / 2;
    for (int i = tid; i < K; i += 512) ssc[i] = shsc[i];
    for (int i = tid; i < K * n1; i += 512) {
        const int k = i / n1;
        const double2 c = coef[(size_t)k * n_c + co1 + i % n1];
        const double s = shsc[k];
        c1b[i] = make_float2((float)(c.x * s), (float)(c.y * s));
    }
    for (int i = tid; i < K * n2; i += 512) {
        const int k = i / n2;
        const double2 c = coef[(size_t)k * n_c + co2 + i % n2];
        const double s = shsc[k];
        c2h[i] = make_uint2(split_pack(c.x * s), split_pack(c.y * s));
    }
    for (int i = tid; i < K * n3; i += 512) {
        const int k = i / n3;
        const double2 c = coef[(size_t)k * n_c + co3 + i % n3];
        const double s = shsc[k];
        c3h[i] = make_uint2(split_pack(c.x * s), split_pack(c.y * s));
    }
    const int e0 = gstart8[goff], ne = gstart8[goff + n3] - e0;
    for (int i = tid; i < ne; i += 512) {
        sw[i] = ew8[e0 + i];
        const uint32_t id = eidx8[e0 + i];
        si[i] = (uint32_t)((int)(id & 0xffffu) - co1) | ((uint32_t)((int)(id >> 16) - co2) << 16);
    }
    for (int i = tid; i <= n3; i += 512) sg[i] = gstart8[goff + i] - e0;
    const int k1lo = 16 * h + 1, k1hi = min(16 * h + 16, K);
    const int off_h = (k1lo - 1) * k1lo / 2;
    const int P = k1hi * (k1hi + 1) / 2 - off_h;
    const int nrb = (P + 15) / 16;
    for (int p = tid; p < nrb * 16; p += 512) {
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
    for (int i = tid; i < Prows * kT16Row; i += 512) Tf[i] = 0.f;
    __syncthreads();

    f32x4 a2[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) a2[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int k1s = 16 * h + l16 + 1;   // stage-1 A row of this lane
    // k1 <= 16 (h = 0): every pair has k2, k3 <= 16, so the second k2 / k3 block is skipped
    const bool two = h > 0;
    const bool okA = k1s <= K, okB0 = l16 + 1 <= K, okB1 = two && l16 + 17 <= K;
    const int nob = two ? 2 * nrb : nrb;   // stage-2 output blocks (pair block, k3 block)
    const int nch = (n3 + 15) / 16;
    for (int ch = 0; ch < nch; ++ch) {
        // ---- stage 1: mu3 groups 16 ch + wave and 16 ch + 8 + wave
        for (int r = 0; r < 2; ++r) {
            const int mu = 16 * ch + wave + 8 * r;
            if (mu >= n3) continue;
            const int s = sg[mu], E8 = sg[mu + 1] - s;   // E8: entries padded to a multiple of 8
            const int nks = (2 * E8 + 31) / 32;
            f32x4 th[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) th[q] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; ks < nks; ++ks) {
                // this lane's 8 entries: K positions 32 ks + 8 lq + j
                const int e8 = 32 * ks + 8 * lq;
                const bool first = e8 < E8;
                const int eb = first ? e8 : e8 - E8;
                u16x8 ah = {}, al = {}, bh[4] = {{}, {}, {}, {}}, bl[4] = {{}, {}, {}, {}};
                if (eb < E8) {
                    const float4 w0 = *reinterpret_cast<const float4 *>(sw + s + eb);
                    const float4 w1 = *reinterpret_cast<const float4 *>(sw + s + eb + 4);
                    const uint4 x0 = *reinterpret_cast<const uint4 *>(si + s + eb);
                    const uint4 x1 = *reinterpret_cast<const uint4 *>(si + s + eb + 4);
                    const float wv[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
                    const uint32_t iv[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int i1 = iv[j] & 0xffff, i2 = iv[j] >> 16;
                        if (okA) {
                            const float2 c = c1b[(k1s - 1) * n1 + i1];
                            const float a = wv[j] * (first ? c.x : c.y);
                            const _Float16 hh = (_Float16)a;
                            ah[j] = __builtin_bit_cast(uint16_t, hh);
                            al[j] = __builtin_bit_cast(uint16_t, (_Float16)(a - (float)hh));
                        }
                        // B' = [Re | -Im] (Tr) and [Im | Re] (Ti) of c_k2, already split
                        if (okB0) {
                            const uint2 c = c2h[l16 * n2 + i2];
                            const uint32_t tr = first ? c.x : (c.y ^ 0x80008000u), ti = first ? c.y : c.x;
                            bh[0][j] = (uint16_t)tr; bl[0][j] = (uint16_t)(tr >> 16);
                            bh[2][j] = (uint16_t)ti; bl[2][j] = (uint16_t)(ti >> 16);
                        }
                        if (okB1) {
                            const uint2 c = c2h[(l16 + 16) * n2 + i2];
                            const uint32_t tr = first ? c.x : (c.y ^ 0x80008000u), ti = first ? c.y : c.x;
                            bh[1][j] = (uint16_t)tr; bl[1][j] = (uint16_t)(tr >> 16);
                            bh[3][j] = (uint16_t)ti; bl[3][j] = (uint16_t)(ti >> 16);
                        }
                    }
                }
                const h16x8 hah = __builtin_bit_cast(h16x8, ah), hal = __builtin_bit_cast(h16x8, al);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if ((q & 1) && !two) continue;
                    const h16x8 hbh = __builtin_bit_cast(h16x8, bh[q]), hbl = __builtin_bit_cast(h16x8, bl[q]);
                    th[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(hal, hbh, th[q], 0, 0, 0);
                    th[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(hah, hbl, th[q], 0, 0, 0);
                    th[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(hah, hbh, th[q], 0, 0, 0);
                }
            }
            // rows k1 = 16h + 4lq + v + 1, columns k2 = 16 (q & 1) + l16 + 1; q >> 1: re / im
            const int sb = 2 * (mu - 16 * ch);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if ((q & 1) && !two) continue;
                const int k2 = 16 * (q & 1) + l16 + 1;
                const int slot = sb + (q >> 1);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k1 = 16 * h + 4 * lq + v + 1;
                    if (k1 <= K && k2 <= k1) {
                        const int p = (k1 - 1) * k1 / 2 + (k2 - 1) - off_h;
                        Tf[p * kT16Row + slot] = th[q][v];
                    }
                }
            }
        }
        __syncthreads();
        // ---- stage 2: As[pair][k3] += T 2^-13 . [Re c_k3 ; -Im c_k3] over this chunk's 32 slots
        {
            // output blocks b = wave + 8 j all have k3 block cb = wave & 1 (0 when h = 0):
            // one B operand per wave
            u16x8 b3h = {}, b3l = {};
            {
                const int k3 = 16 * (two ? (wave & 1) : 0) + l16 + 1;
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) {
                    const int m3 = 16 * ch + 4 * lq + (jj >> 1);
                    if (m3 < n3 && k3 <= K) {
                        const uint2 c = c3h[(k3 - 1) * n3 + m3];
                        const uint32_t v = (jj & 1) ? (c.y ^ 0x80008000u) : c.x;
                        b3h[jj] = (uint16_t)v;
                        b3l[jj] = (uint16_t)(v >> 16);
                    }
                }
            }
            const h16x8 bh = __builtin_bit_cast(h16x8, b3h), bl = __builtin_bit_cast(h16x8, b3l);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int b = wave + 8 * j;
                if (b < nob) {
                    const int rb = two ? b >> 1 : b;
                    const float4 *tp = reinterpret_cast<const float4 *>(Tf + (16 * rb + l16) * kT16Row + 8 * lq);
                    const float4 t0 = tp[0], t1 = tp[1];
                    const float tv[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
                    h16x8 ah, al;
#pragma unroll
                    for (int jj = 0; jj < 8; ++jj) split_h(tv[jj] * (1.0f / 8192.0f), jj, ah, al);
                    a2[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, a2[j], 0, 0, 0);
                    a2[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, a2[j], 0, 0, 0);
                    a2[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, a2[j], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    // rows = pairs p = 16 rb + 4 lq + v, columns k3 = 16 cb + l16 + 1
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int b = wave + 8 * j;
        if (b < nob) {
            const int rb = two ? b >> 1 : b, cb = two ? b & 1 : 0;
            const int k3 = 16 * cb + l16 + 1;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int p = 16 * rb + 4 * lq + v;
                if (p < P) {
                    const int k1 = pk[p] & 0xff, k2 = pk[p] >> 8;
                    if (k3 <= k2) {
                        const int64_t kt = (int64_t)(k1 - 1) * k1 * (k1 + 1) / 6 + (int64_t)(k2 - 1) * k2 / 2 + (k3 - 1);
                        const double val = (double)a2[j][v] * 8192.0 / (ssc[k1 - 1] * ssc[k2 - 1] * ssc[k3 - 1]);
                        As[(int64_t)lt * ldA + as_col(kt, cbase, nprop, p0)] = val;
                    }
                }
            }
        }
    }
}

__global__ __launch_bounds__(512) void k_triple16(const double2 *__restrict__ cb0, const double2 *__restrict__ cb1,
                                                  const int *__restrict__ sel, int flip, const double *__restrict__ shsc,
                                                  int n_c, int K, const float *__restrict__ ew,
                                                  const uint32_t *__restrict__ eidx, const int *__restrict__ gstart,
                                                  const int4 *__restrict__ ltinfo, const int *__restrict__ work, int Lw,
                                                  int Emax, int Prows, double *__restrict__ As, int64_t kl_pad)
{
    extern __shared__ __align__(16) char smem16[];
    const int wk = work[blockIdx.x];
    triple16_body(smem16, pick(cb0, cb1, sel, flip), shsc, n_c, K, wk & 0xffff, wk >> 16, ew, eidx, gstart, ltinfo,
                  Lw, Emax, Prows, As, kl_pad, 0, 1, 0);
}

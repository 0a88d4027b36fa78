// tp_mlp.hpp -- gfx950 device code of the learned rotation proposer (included
// by tp_hip.hip after tp_batch.hpp).
//
// Network: 192 inputs (190 real-basis coefficients of one shell, normalised,
// log10 T, one zero pad) -> 256 -> 256 -> 256 (ReLU) -> 64 outputs (the first
// 36 are one quaternion per even l = 2..18; the rest are padding rows whose
// weights stay zero).  BF16 operands, FP32 accumulation on
// v_mfma_f32_16x16x32_bf16, FP32 master weights and Adam moments.
#pragma once

#include "tp_batch.hpp"

constexpr int kMlpIn = 192, kMlpHid = 256, kMlpOut = 64, kMlpQ = 36, kMlpFeat = 190;
// flat parameter layout: W1 [256][192], b1, W2 [256][256], b2, W3 [256][256], b3, W4 [64][256], b4
constexpr int kMlpOffW[4] = {0, 256 * 192 + 256, 256 * 192 + 256 + 256 * 256 + 256,
                             256 * 192 + 256 + 2 * (256 * 256 + 256)};
constexpr int kMlpRows[4] = {kMlpHid, kMlpHid, kMlpHid, kMlpOut};
constexpr int kMlpCols[4] = {kMlpIn, kMlpHid, kMlpHid, kMlpHid};
constexpr int kMlpNParam = 256 * 192 + 256 + 2 * (256 * 256 + 256) + 64 * 256 + 64;

__device__ __forceinline__ float bf16_to_f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t f_to_bf16(float f)   // round to nearest even
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------------------
// C[M][N] = scale * sum_k A[M][K] B[N][K] (+ bias[N]) (ReLU) (* [mask > 0]),
// A and B BF16 with K contiguous; 64 x 64 tile per workgroup of 4 waves (2x2,
// 32 x 32 each), K staged 32 at a time.  M, N multiples of 64, K of 32.
// Outputs (any subset): FP32 [M][N], BF16 [M][N], BF16 transposed [N][M].
// ---------------------------------------------------------------------------
struct MlpEpi {
    const float *bias;
    int relu;
    float scale;
    const uint16_t *mask;
    int64_t ldm;
    float *outf;
    int64_t ldf;
    uint16_t *outb;
    int64_t ldo;
    uint16_t *outt;
    int64_t ldt;
};

__global__ __launch_bounds__(256) void k_mlp_gemm(const uint16_t *__restrict__ A, int64_t lda,
                                                  const uint16_t *__restrict__ B, int64_t ldb, int M, int N, int K,
                                                  MlpEpi ep)
{
    __shared__ uint16_t sA[64][40];
    __shared__ uint16_t sB[64][40];
    const int ntn = N / 64;
    const int r0 = (blockIdx.x / ntn) * 64, c0 = (blockIdx.x % ntn) * 64;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1, l16 = lane & 15, lq = lane >> 4;
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int row = tid >> 2, kc = (tid & 3) * 8;
    for (int k0 = 0; k0 < K; k0 += 32) {
        const u16x8 va = *reinterpret_cast<const u16x8 *>(A + (int64_t)(r0 + row) * lda + k0 + kc);
        const u16x8 vb = *reinterpret_cast<const u16x8 *>(B + (int64_t)(c0 + row) * ldb + k0 + kc);
        __syncthreads();
        *reinterpret_cast<u16x8 *>(&sA[row][kc]) = va;
        *reinterpret_cast<u16x8 *>(&sB[row][kc]) = vb;
        __syncthreads();
        b16x8 fa[2], fb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            fa[i] = __builtin_bit_cast(b16x8, *reinterpret_cast<const u16x8 *>(&sA[wm * 32 + i * 16 + l16][8 * lq]));
            fb[i] = __builtin_bit_cast(b16x8, *reinterpret_cast<const u16x8 *>(&sB[wn * 32 + i * 16 + l16][8 * lq]));
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = c0 + wn * 32 + j * 16 + l16;
            const int rb = r0 + wm * 32 + i * 16 + 4 * lq;
            const float bs = ep.bias ? ep.bias[col] : 0.f;
            uint16_t tv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rb + v;
                float x = acc[i][j][v] * ep.scale + bs;
                if (ep.relu) x = fmaxf(x, 0.f);
                if (ep.mask && !(bf16_to_f(ep.mask[(int64_t)r * ep.ldm + col]) > 0.f)) x = 0.f;
                if (ep.outf) ep.outf[(int64_t)r * ep.ldf + col] = x;
                const uint16_t h = f_to_bf16(x);
                if (ep.outb) ep.outb[(int64_t)r * ep.ldo + col] = h;
                tv[v] = h;
            }
            if (ep.outt) {
                uint16_t *t = ep.outt + (int64_t)col * ep.ldt + rb;
                t[0] = tv[0]; t[1] = tv[1]; t[2] = tv[2]; t[3] = tv[3];
            }
        }
}

// Sum over rows of a transposed BF16 matrix: out[n] = sum_b Xt[n][b] (bias
// gradients), one wave per n, fixed order.
__global__ __launch_bounds__(256) void k_mlp_rowsum(const uint16_t *__restrict__ Xt, int64_t ld, int cols,
                                                    float *__restrict__ out, int n)
{
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (w >= n) return;
    float s = 0.f;
    for (int b = lane; b < cols; b += 64) s += bf16_to_f(Xt[(int64_t)w * ld + b]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[w] = s;
}

// Loss and output gradient: L = (1/B) sum_b sum_{j<36} (y - t)^2;
// dY = 2 (y - t) / B as BF16 in [B][64] and transposed [64][B] (zero past 36).
// Per-row partial losses go to lpart[B] (the host or a later kernel sums them).
__global__ __launch_bounds__(256) void k_mlp_loss(const float *__restrict__ Y, const float *__restrict__ T, int Bn,
                                                  uint16_t *__restrict__ dY, uint16_t *__restrict__ dYt,
                                                  float *__restrict__ lpart)
{
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6), j = threadIdx.x & 63;
    if (b >= Bn) return;
    float d = 0.f, l = 0.f;
    if (j < kMlpQ) {
        const float e = Y[(int64_t)b * kMlpOut + j] - T[(int64_t)b * kMlpQ + j];
        l = e * e;
        d = 2.f * e / (float)Bn;
    }
    const uint16_t h = f_to_bf16(d);
    dY[(int64_t)b * kMlpOut + j] = h;
    dYt[(int64_t)j * Bn + b] = h;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if (j == 0) lpart[b] = l / (float)Bn;
}

// Adam on the FP32 master parameters; refreshes the BF16 copies W [rows][cols]
// and W^T [cols][rows] of the weight matrices.  t = step count (1-based) read
// from the device so that a captured graph can be replayed.
struct MlpWeights {
    uint16_t *W[4];
    uint16_t *Wt[4];
};

__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                                  const float *__restrict__ g, const int *__restrict__ tstep, float lr,
                                                  float b1, float b2, float eps, MlpWeights w)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kMlpNParam) return;
    const int t = tstep[0] + 1;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.f - powf(b1, (float)t)), vh = vi / (1.f - powf(b2, (float)t));
    float x = p[i] - lr * mh / (sqrtf(vh) + eps);
    // weight matrix entries: keep the padding rows of the output layer at zero
#pragma unroll
    for (int L = 0; L < 4; ++L) {
        const int o = kMlpOffW[L], n = kMlpRows[L] * kMlpCols[L];
        if (i >= o && i < o + n) {
            const int r = (i - o) / kMlpCols[L], c = (i - o) % kMlpCols[L];
            if (L == 3 && r >= kMlpQ) x = 0.f;
            const uint16_t h = f_to_bf16(x);
            w.W[L][(int64_t)r * kMlpCols[L] + c] = h;
            w.Wt[L][(int64_t)c * kMlpRows[L] + r] = h;
        }
        if (L == 3 && i >= o + n + kMlpQ) x = 0.f;   // output bias padding
    }
    p[i] = x;
}

__global__ void k_mlp_tick(int *tstep) { tstep[0] += 1; }

// ---------------------------------------------------------------------------
// Features of one shell: real-basis coefficients x (x_0 = Re c_0,
// x_m = sqrt2 s Re c_m, x_-m = -sqrt2 s Im c_m, s = (-1)^m, m > 0; the real
// form of a real function), scaled to norm 8, then log10 T and a zero.
// ---------------------------------------------------------------------------
__device__ inline void mlp_features(const double2 *__restrict__ c, int n_c, double T, uint16_t *__restrict__ xb,
                                    uint16_t *__restrict__ xt, int64_t ldt, int row, float *__restrict__ xf, int lane)
{
    // lane-strided over the (at most 190) compact entries of l <= 18; the norm
    // by a wave reduction; entries past n_c (L < 18) are zero
    const int nf = n_c < kMlpFeat ? n_c : kMlpFeat;
    double ss = 0.0;
    for (int i = lane; i < nf; i += 64) { const double2 v = c[i]; ss += v.x * v.x + v.y * v.y; }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
    const double sc = ss > 0.0 ? 8.0 / sqrt(ss) : 0.0;
    const double r2 = 1.4142135623730951;
    for (int i = lane; i < kMlpIn; i += 64) {
        double x = 0.0;
        if (i < nf) {
            int l = 0;
            while ((l + 2) * (l + 1) / 2 <= i) l += 2;   // block of entry i: offset l(l-1)/2
            const int m = i - l * (l - 1) / 2 - l;
            const int am = m < 0 ? -m : m;
            const double s = (am & 1) ? -1.0 : 1.0;
            const double2 cm = c[l * (l - 1) / 2 + l + am];
            x = m == 0 ? cm.x : (m > 0 ? r2 * s * cm.x : -r2 * s * cm.y);
            x *= sc;
        } else if (i == kMlpFeat) {
            x = T > 0.0 ? log10(T) : -30.0;
        }
        const uint16_t h = f_to_bf16((float)x);
        if (xb) xb[(int64_t)row * kMlpIn + i] = h;
        if (xt) xt[(int64_t)i * ldt + row] = h;
        if (xf) xf[(int64_t)row * kMlpIn + i] = (float)x;
    }
}

// Inference input of the learned slots: row g * nlearn + j holds shell
// j * K_g / nlearn of run g's current state and the run's temperature.
__global__ __launch_bounds__(64) void k_mlp_infer_features(const GroupState *__restrict__ gs,
                                                           const double2 *__restrict__ cur, int64_t cstride, int n_c,
                                                           int nlearn, uint16_t *__restrict__ X)
{
    const int g = blockIdx.y, j = blockIdx.x;
    const GroupState s = gs[g];
    const int k = j * s.K / nlearn;
    mlp_features(cur + (int64_t)g * cstride + (int64_t)k * n_c, n_c, s.T, X, nullptr, 0, g * nlearn + j, nullptr,
                 threadIdx.x);
}

__device__ inline double4 qmul(double4 a, double4 b)   // Hamilton product (w, x, y, z)
{
    return make_double4(a.x * b.x - a.y * b.y - a.z * b.z - a.w * b.w, a.x * b.y + a.y * b.x + a.z * b.w - a.w * b.z,
                        a.x * b.z - a.y * b.w + a.z * b.x + a.w * b.y, a.x * b.w + a.y * b.z - a.z * b.y + a.w * b.x);
}

// Network outputs -> quaternions of the learned slots, each composed with a
// random rotation of angle <= noise * amax * pi/4 (exploration).
__global__ __launch_bounds__(256) void k_mlp_to_quat(const EvalDesc *__restrict__ ev, int ne, int nl,
                                                     const GroupState *__restrict__ gs, const float *__restrict__ Y,
                                                     double noise, double *__restrict__ quat)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne * nl) return;
    const int e = i / nl, li = i % nl;
    const EvalDesc d = ev[e];
    const GroupState g = gs[d.group];
    const int j = d.p - (d.nprop - g.nlearn);
    if (j < 0) return;
    const float *y = Y + (int64_t)(d.group * g.nlearn + j) * kMlpOut + 4 * li;
    double4 q = make_double4(y[0], y[1], y[2], y[3]);
    const double n = sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (n > 1e-6) q = make_double4(q.x / n, q.y / n, q.z / n, q.w / n);
    else q = make_double4(1.0, 0.0, 0.0, 0.0);
    if (noise > 0.0) {
        double u0, u1, u2, u3, u4, u5;
        philox_u01x2(g.seed, (uint32_t)li, (uint32_t)d.p, (uint32_t)g.nsweep, TP_TAG_QAXIS, u0, u1);
        philox_u01x2(g.seed, (uint32_t)li, (uint32_t)d.p, (uint32_t)g.nsweep, TP_TAG_QAXIS + 1u, u2, u3);
        philox_u01x2(g.seed, (uint32_t)li, (uint32_t)d.p, (uint32_t)g.nsweep, TP_TAG_QANGLE, u4, u5);
        const double r0 = sqrt(-2.0 * log(1.0 - u0)), r1 = sqrt(-2.0 * log(1.0 - u2));
        double ax = r0 * cos(2.0 * kPiD * u1), ay = r0 * sin(2.0 * kPiD * u1), az = r1 * cos(2.0 * kPiD * u3);
        const double an = sqrt(ax * ax + ay * ay + az * az);
        if (an > 0.0) { ax /= an; ay /= an; az /= an; } else { ax = 0.0; ay = 0.0; az = 1.0; }
        const double th = haar_angle(u4, noise * g.amax * 0.25 * kPiD);
        const double sh = sin(0.5 * th);
        q = qmul(make_double4(cos(0.5 * th), sh * ax, sh * ay, sh * az), q);
    }
    double *o = quat + (size_t)i * 4;
    o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
}

// ---------------------------------------------------------------------------
// Move log: after a sweep's evaluations, proposals that pass the Metropolis
// test on their own against the run's current energy (dE < 0, or u < exp(-dE/T)
// with u from the run's stream, tag LOG) are logged, one sample per shell of
// the run: features of that shell of the state the move starts from, and the
// proposal's quaternions (w >= 0 form; for plane moves the angle quaternions
// that k_rotate_batch writes back).  One workgroup walks all proposals in
// order, so the log is deterministic.  Runs before k_select_batch.
// ---------------------------------------------------------------------------
#define TP_TAG_LOG 0x4C4F4721u

__global__ __launch_bounds__(1024) void k_log_moves(const EvalDesc *__restrict__ ev, int ne, int nl,
                                                    const GroupState *__restrict__ gs, const double *__restrict__ Eprop,
                                                    const double *__restrict__ quat, int n_c,
                                                    uint16_t *__restrict__ Xlog, float *__restrict__ Tlog,
                                                    int *__restrict__ count, int cap)
{
    __shared__ int s_off[1024];
    __shared__ int s_base;
    const int tid = threadIdx.x;
    if (tid == 0) s_base = count[0];
    __syncthreads();
    for (int e0 = 0; e0 < ne; e0 += 1024) {
        const int e = e0 + tid;
        int take = 0;
        if (e < ne) {
            const EvalDesc d = ev[e];
            const GroupState g = gs[d.group];
            const double dE = Eprop[e] - g.E;
            double u0, u1;
            philox_u01x2(g.seed, (uint32_t)d.p, 0u, (uint32_t)g.nsweep, TP_TAG_LOG, u0, u1);
            const bool ok = dE < 0.0 || (dE > 0.0 && g.T > 0.0 && u0 < exp(-dE / g.T));
            take = ok ? d.K : 0;
        }
        s_off[tid] = take;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {   // inclusive scan
            const int v = tid >= o ? s_off[tid - o] : 0;
            __syncthreads();
            s_off[tid] += v;
            __syncthreads();
        }
        const int base = s_base;
        const int total = s_off[1023];
        __syncthreads();
        // the samples: wave w handles samples w, w + 16, ... of this chunk
        const int wave = tid >> 6, lane = tid & 63;
        for (int sidx = wave; sidx < total; sidx += 16) {
            // find the proposal of sample sidx (binary search in the scan)
            int lo = 0, hi = 1023;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_off[mid] > sidx) hi = mid; else lo = mid + 1;
            }
            const int ee = e0 + lo;
            const int first = lo ? s_off[lo - 1] : 0;
            const int k = sidx - first;
            const int slot = base + sidx;
            if (slot >= cap) break;
            const EvalDesc d = ev[ee];
            const GroupState g = gs[d.group];
            mlp_features(d.src + (int64_t)k * n_c, n_c, g.T, Xlog, nullptr, 0, slot, nullptr, lane);
            if (lane < 4 * nl && lane < kMlpQ) {
                const int li = lane >> 2, c = lane & 3;
                const double *q = quat + ((size_t)ee * nl + li) * 4;
                const double sg = q[0] < 0.0 ? -1.0 : 1.0;
                Tlog[(int64_t)slot * kMlpQ + lane] = (float)(sg * q[c]);
            }
        }
        __syncthreads();
        if (tid == 0) s_base = base + total;
        __syncthreads();
    }
    if (tid == 0) count[0] = s_base < cap ? s_base : cap;
}

// Minibatch gather for training: rows idx[b] of the log -> X [B][192] and
// X^T [192][B] (BF16), targets [B][36] (FP32).
__global__ __launch_bounds__(256) void k_mlp_gather(const uint16_t *__restrict__ Xlog, const float *__restrict__ Tlog,
                                                    const int *__restrict__ count, uint64_t seed,
                                                    const int *__restrict__ tstep, int Bn, uint16_t *__restrict__ X,
                                                    uint16_t *__restrict__ Xt, float *__restrict__ T)
{
    const int b = blockIdx.x;
    const int n = count[0];
    if (n <= 0) return;
    uint32_t c[4] = {(uint32_t)b, (uint32_t)tstep[0], 0u, TP_TAG_LOG + 1u};
    philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const int r = (int)(((uint64_t)c[0] * (uint64_t)n) >> 32);
    for (int i = threadIdx.x; i < kMlpIn; i += 256) {
        const uint16_t h = Xlog[(int64_t)r * kMlpIn + i];
        X[(int64_t)b * kMlpIn + i] = h;
        Xt[(int64_t)i * Bn + b] = h;
    }
    for (int i = threadIdx.x; i < kMlpQ; i += 256) T[(int64_t)b * kMlpQ + i] = Tlog[(int64_t)r * kMlpQ + i];
}

// tp_grad.hpp -- gradient of the Bayes energy with respect to the shell
// coefficients, FP64 (the chain rule through energy(), correlations.jl:201-226,
// the contraction c3raw = Bm.(H.As) and the triple products, :165-174; the
// reference evaluates the energy only).
//
//   D[r,kt]   = dE/dc3[r,kt] = tF (qsum / S - qhat_r / c3_r)  where c3_r > 1e-30
//               (0 where the clamp max(c3, 1e-30) is active), S = sum_r max(c3_r, 1e-30)
//   gG        = Bm^T D   (k_c3sep_bwd: the separated grid run backwards)
//   gAs       = H^T gG   (FP64 MFMA GEMM, k_gemm; on the curved Ewald sphere
//                         H depends on the k-triple: k_glam with the terms
//                         grouped by l-triple, after the row blocks of gG are summed)
//   gamma_k   = dE/dRe c_k + i dE/dIm c_k  (k_triple_bwd + k_grad_reduce)
//
// With As[lt,kt] = Re sum_mu3 T_{k1k2}[lt,mu3] c_k3[l3,mu3] and
// T_{k1k2}[lt,mu3] = sum_{e in (lt,mu3)} w_e c_k1[i1_e] c_k2[i2_e], and
// A_{k1k2}[lt,mu3] = sum_{k3 <= k2} gAs[lt,kt(k1,k2,k3)] c_k3[l3,mu3]:
//   slot 3: gamma_k3[l3,mu3] += sum_{k1 >= k2 >= k3} gAs[lt,kt] conj(T_{k1k2}[lt,mu3])
//   slot 1: gamma_k1[i1_e]   += w_e conj(sum_{k2 <= k1} A_{k1k2}[lt,mu3] c_k2[i2_e])
//   slot 2: gamma_k2[i2_e]   += w_e conj(sum_{k1 >= k2} A_{k1k2}[lt,mu3] c_k1[i1_e])
// (d Re(z w) = Re(conj(conj(w)) dz): the complex gradient of Re(z w) in z is conj(w)).
#pragma once

// ---------------------------------------------------------------------------
// Backward of the separated-grid contraction with D formed on the fly.
// Workgroup = (16 k-triple columns, 16 values of j) as in k_c3sep.  Stage A
// (CS = TA.G) and stage B (c3 = TB.CS, here in the orientation rows i x
// columns col, so that the accumulator of a block holds rows i = lq + 4v)
// recompute the workgroup's c3 tile; D replaces it in the registers; then
//   PS[k'][jl][col] = sum_i TB[i][k'] D[i][jl][col]
// runs on FP64 MFMA with D as the B operand straight from the accumulators
// (k-step s of a 16-row block takes rows 4s + lq: the accumulator's v = s),
// and gGp[jc][cls][col] = sum_{jl,cs} TA[cs][cls][j0 + jl] PS[2 m2i(cls) + cs][jl][col]
// on the VALU.  Waves own whole values of j (both i blocks), so PS needs no
// cross-wave sum.  The partial over this workgroup's 16 values of j goes to
// row block jc of gGp; the GEMM that follows sums the njc blocks.
// LDS: CS / PS [kb][16][17] doubles (dynamic, c3sep_lds_bytes(kb)).
// QB = kSepWideQB: the wide layout of k_c3sep (odd-m classes of the curved
// Ewald sphere: classes per m2 in chunks of 24, kb <= 40, three 16-row
// blocks of PS).
// ---------------------------------------------------------------------------
template <int QB = 6>
__global__ __launch_bounds__(512) void k_c3sep_bwd(const double *__restrict__ G, int64_t ldg,
                                                   const int *__restrict__ m2start, int n_m2,
                                                   const double *__restrict__ TA, int64_t ta_plane, int jpad,
                                                   const double *__restrict__ TB, int kb, int N, int ncols,
                                                   const double *__restrict__ qhat, const double *__restrict__ qsum,
                                                   const double *__restrict__ tf, const double *__restrict__ part,
                                                   int64_t ldp, int nprt, int n_m, double *__restrict__ gGp,
                                                   int64_t ldgg, int nm_pad)
{
    extern __shared__ double cs_dyn[];
    auto CS = [&](int k, int r, int c) -> double & { return cs_dyn[(k * 16 + r) * 17 + c]; };
    constexpr int NW = 8;
    const int c0 = blockIdx.x * 16, jc = blockIdx.y, j0 = jc * 16;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int l16 = lane & 15, lq = lane >> 4;
    const int R = 2 * N * N, twoN = 2 * N;
    const int nib = (N + 15) / 16;
    for (int i = tid; i < (kb - 2 * n_m2) * 256; i += 512) {
        const int k = 2 * n_m2 + i / 256, r = (i / 16) % 16, c = i % 16;
        CS(k, r, c) = 0.0;
    }
    // ---- forward stage A (as k_c3sep)
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
    // per-column energy terms of this lane's column (col = c0 + l16)
    const int col = c0 + l16;
    const bool cok = col < ncols;
    const int colc = min(col, ncols - 1);
    double S = 0.0;
    for (int q = 0; q < nprt; ++q) S += part[(int64_t)(2 * q) * ldp + colc];   // the forward's nprt partial rows
    const double tfc = tf[colc], qsc = qsum[colc];
    __syncthreads();
    // ---- forward stage B in the orientation (i, col) and D; this wave's
    // values of j: jl = wave and wave + 8, both i blocks
    dbl4 dv[2][2];   // [jl slot][ib]
#pragma unroll
    for (int js = 0; js < 2; ++js) {
        const int jl = wave + NW * js, j = j0 + jl;
#pragma unroll
        for (int ib = 0; ib < 2; ++ib) {
            dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};
            if (ib < nib) {
#pragma unroll
                for (int q = 0; q < QB; ++q)
                    if (4 * q < kb) {
                        const int kk = 4 * q + lq;
                        const double a = TB[(int64_t)(16 * ib + l16) * kb + min(kk, kb - 1)];
                        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(kk < kb ? a : 0.0, CS(kk < kb ? kk : 0, jl, l16),
                                                                   acc, 0, 0, 0);
                    }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 16 * ib + lq + 4 * v;
                double d = 0.0;
                if (ib < nib && i < N && j < twoN && cok) {
                    const double c = acc[v];
                    if (c > 1e-30) d = tfc * (qsc / S - qhat[(int64_t)col * R + i + N * j] / c);
                }
                dv[js][ib][v] = d;
            }
        }
    }
    __syncthreads();   // CS is read by every wave above; PS reuses its storage
    // ---- PS[k'][jl][col] = sum_i TB[i][k'] D[i][jl][col]: rows k' (two blocks of 16), columns col
#pragma unroll
    for (int js = 0; js < 2; ++js) {
        const int jl = wave + NW * js;
#pragma unroll
        for (int kblk = 0; kblk < (4 * QB + 15) / 16; ++kblk) {
            dbl4 acc = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int ib = 0; ib < 2; ++ib) {
                if (ib >= nib) continue;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    const int i = 16 * ib + 4 * s + lq, kp = 16 * kblk + l16;
                    const double a = (i < N && kp < kb) ? TB[(int64_t)i * kb + kp] : 0.0;
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, dv[js][ib][s], acc, 0, 0, 0);
                }
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int kp = 16 * kblk + lq + 4 * v;
                if (kp < kb) CS(kp, jl, l16) = acc[v];
            }
        }
    }
    __syncthreads();
    // ---- gGp[jc][cls][col] = sum_{jl, cs} TA[cs][cls][j0 + jl] PS[2 m2i + cs][jl][col]
    for (int it = tid; it < n_m * 16; it += 512) {
        const int cls = it >> 4, cl = it & 15;
        int m2i = 0;
        while (m2i + 1 < n_m2 && m2start[m2i + 1] <= cls) ++m2i;
        double s = 0.0;
#pragma unroll
        for (int cs = 0; cs < 2; ++cs) {
            const double *ta = TA + (int64_t)cs * ta_plane + (int64_t)cls * jpad + j0;
#pragma unroll 4
            for (int jl = 0; jl < 16; ++jl)
                if (j0 + jl < twoN) s += ta[jl] * CS(2 * m2i + cs, jl, cl);
        }
        gGp[((int64_t)jc * nm_pad + cls) * ldgg + c0 + cl] = s;
    }
}

// ---------------------------------------------------------------------------
// Backward of the triple products, one workgroup per l-triple (1024
// threads, FP64 VALU).  LDS: the K shells' l1, l2, l3 blocks, the l-triple's
// row of gAs, its entries, per-pair T and A of the current mu3, the three
// gradient blocks and the slot-3 partial sums.  Per mu3: phase 1 (one thread
// per pair k1 >= k2) forms T and A; phase 2 accumulates slots 1 and 2 (one
// thread per (entry, k); within a group the entries have distinct m1 and
// distinct m2, so no two threads write the same element) and the slot-3
// partials over 16 fixed chunks of pairs (one thread per (chunk, k3)); the
// chunks are summed in fixed order by the first threads of the next phase 1.
// Output: gpart[lt][slot][k][m] (m = 0..2l+1 of the slot's l, ld Lw).
// ---------------------------------------------------------------------------
constexpr int kBwdChunks = 16;

inline size_t triple_bwd_lds_bytes(int K, int Lw, int64_t klength, int Emax)
{
    const size_t np = (size_t)K * (K + 1) / 2;
    return sizeof(double2) * (6 * (size_t)K * Lw + 2 * np + (size_t)kBwdChunks * 32) + sizeof(double) * (klength + Emax) +
           sizeof(uint32_t) * Emax + sizeof(int) * (Lw + 1 + np + 1);
}

__global__ __launch_bounds__(1024) void k_triple_bwd(const double2 *__restrict__ coef, int n_c, int K,
                                                    const double *__restrict__ ew, const uint32_t *__restrict__ eidx,
                                                    const int *__restrict__ gstart, const int4 *__restrict__ ltinfo,
                                                    int Lw, int Emax, const double *__restrict__ gA, int64_t ldga,
                                                    int klength, const int *__restrict__ order,
                                                    double2 *__restrict__ gpart)
{
    extern __shared__ double2 smb[];
    const int lt = order[blockIdx.x];   // most expensive l-triples first
    const int4 li = ltinfo[lt];
    const int l3 = li.z, l1 = li.w & 0xff, l2 = (li.w >> 8) & 0xff;
    const int n1 = 2 * l1 + 1, n2 = 2 * l2 + 1, n3 = 2 * l3 + 1;
    const int goff = li.y;
    const int np = K * (K + 1) / 2;
    double2 *c1b = smb;                          // [K][Lw]
    double2 *c2b = c1b + (size_t)K * Lw;
    double2 *c3b = c2b + (size_t)K * Lw;
    double2 *g1 = c3b + (size_t)K * Lw;          // gradient blocks [K][Lw]
    double2 *g2 = g1 + (size_t)K * Lw;
    double2 *g3 = g2 + (size_t)K * Lw;
    double2 *Ts = g3 + (size_t)K * Lw;           // [np]
    double2 *Ap = Ts + np;                       // [np]
    double2 *g3p = Ap + np;                      // [kBwdChunks][32]
    double *sgA = reinterpret_cast<double *>(g3p + kBwdChunks * 32);   // [klength]
    double *sw = sgA + klength;                  // [Emax]
    uint32_t *si = reinterpret_cast<uint32_t *>(sw + Emax);            // [Emax]
    int *sg = reinterpret_cast<int *>(si + Emax);                      // [Lw + 1]
    int *pk = sg + (Lw + 1);                                           // [np] k1 | k2 << 8 (1-based)

    const int tid = threadIdx.x;
    constexpr int NT = 1024;
    const int npr = (np + 63) & ~63;   // first thread of the A group (2 npr <= NT - 32: K <= 27)
    const int co1 = l1 * (l1 - 1) / 2, co2 = l2 * (l2 - 1) / 2, co3 = l3 * (l3 - 1) / 2;
    for (int i = tid; i < K * n1; i += NT) c1b[(i / n1) * Lw + i % n1] = coef[(size_t)(i / n1) * n_c + co1 + i % n1];
    for (int i = tid; i < K * n2; i += NT) c2b[(i / n2) * Lw + i % n2] = coef[(size_t)(i / n2) * n_c + co2 + i % n2];
    for (int i = tid; i < K * n3; i += NT) c3b[(i / n3) * Lw + i % n3] = coef[(size_t)(i / n3) * n_c + co3 + i % n3];
    for (int i = tid; i < 3 * K * Lw; i += NT) g1[i] = make_double2(0.0, 0.0);
    for (int i = tid; i < klength; i += NT) sgA[i] = gA[(int64_t)lt * ldga + i];
    const int e0 = gstart[goff], ne = gstart[goff + n3] - e0;
    for (int i = tid; i < ne; i += NT) {
        sw[i] = ew[e0 + i];
        const uint32_t id = eidx[e0 + i];
        si[i] = (uint32_t)((int)(id & 0xffffu) - co1) | ((uint32_t)((int)(id >> 16) - co2) << 16);
    }
    for (int i = tid; i <= n3; i += NT) sg[i] = gstart[goff + i] - e0;
    for (int p = tid; p < np; p += NT) {
        int k1 = (int)((1.0 + sqrt(1.0 + 8.0 * p)) * 0.5);
        while (k1 * (k1 - 1) / 2 > p) --k1;
        while ((k1 + 1) * k1 / 2 <= p) ++k1;
        pk[p] = k1 | ((p - k1 * (k1 - 1) / 2 + 1) << 8);
    }
    __syncthreads();

    auto ktof = [](int k1, int k2, int k3) { return (k1 - 1) * k1 * (k1 + 1) / 6 + (k2 - 1) * k2 / 2 + (k3 - 1); };
    for (int mu = 0; mu <= n3; ++mu) {
        // slot-3 partials of the previous mu3, summed in chunk order
        if (mu > 0 && tid >= NT - 32 && tid - (NT - 32) < K) {
            const int k3 = tid - (NT - 32);
            double2 s = make_double2(0.0, 0.0);
            for (int ch = 0; ch < kBwdChunks; ++ch) {
                s.x += g3p[ch * 32 + k3].x;
                s.y += g3p[ch * 32 + k3].y;
            }
            g3[k3 * Lw + mu - 1].x += s.x;
            g3[k3 * Lw + mu - 1].y += s.y;
        }
        if (mu == n3) break;
        const int s0 = sg[mu], E = sg[mu + 1] - s0;
        // phase 1: T of every pair (threads [0, np)) and A of every pair
        // (threads [npr, npr + np)); two accumulators per sum for ILP
        if (tid < np) {
            const int p = tid;
            const int k1 = pk[p] & 0xff, k2 = pk[p] >> 8;
            const double2 *a = c1b + (k1 - 1) * Lw, *b = c2b + (k2 - 1) * Lw;
            double tr0 = 0.0, ti0 = 0.0, tr1 = 0.0, ti1 = 0.0;
            int e = 0;
            for (; e + 1 < E; e += 2) {
                const uint32_t id0 = si[s0 + e], id1 = si[s0 + e + 1];
                const double w0 = sw[s0 + e], w1 = sw[s0 + e + 1];
                const double2 x0 = a[id0 & 0xffffu], y0 = b[id0 >> 16], x1 = a[id1 & 0xffffu], y1 = b[id1 >> 16];
                tr0 += w0 * (x0.x * y0.x - x0.y * y0.y);
                ti0 += w0 * (x0.x * y0.y + x0.y * y0.x);
                tr1 += w1 * (x1.x * y1.x - x1.y * y1.y);
                ti1 += w1 * (x1.x * y1.y + x1.y * y1.x);
            }
            if (e < E) {
                const uint32_t id0 = si[s0 + e];
                const double w0 = sw[s0 + e];
                const double2 x0 = a[id0 & 0xffffu], y0 = b[id0 >> 16];
                tr0 += w0 * (x0.x * y0.x - x0.y * y0.y);
                ti0 += w0 * (x0.x * y0.y + x0.y * y0.x);
            }
            Ts[p] = make_double2(tr0 + tr1, ti0 + ti1);
        } else if (tid >= npr && tid < npr + np) {
            const int p = tid - npr;
            const int k1 = pk[p] & 0xff, k2 = pk[p] >> 8;
            double ar0 = 0.0, ai0 = 0.0, ar1 = 0.0, ai1 = 0.0;
            const int kb = ktof(k1, k2, 1);
            int k3 = 1;
            for (; k3 + 1 <= k2; k3 += 2) {
                const double g0 = sgA[kb + k3 - 1], g1v = sgA[kb + k3];
                const double2 c0 = c3b[(k3 - 1) * Lw + mu], c1 = c3b[k3 * Lw + mu];
                ar0 += g0 * c0.x;
                ai0 += g0 * c0.y;
                ar1 += g1v * c1.x;
                ai1 += g1v * c1.y;
            }
            if (k3 <= k2) {
                const double g0 = sgA[kb + k3 - 1];
                const double2 c0 = c3b[(k3 - 1) * Lw + mu];
                ar0 += g0 * c0.x;
                ai0 += g0 * c0.y;
            }
            Ap[p] = make_double2(ar0 + ar1, ai0 + ai1);
        }
        __syncthreads();
        // phase 2a: slots 1 and 2, thread (entry e, shell k)
        if (tid < 512) {
            for (int t = tid; t < E * K; t += 512) {
                const int e = t / K, k = t % K + 1;
                const uint32_t id = si[s0 + e];
                const int i1 = id & 0xffffu, i2 = id >> 16;
                const double w = sw[s0 + e];
                // slot 1 (k1 = k): sum_{k2 <= k} A_{k,k2} c_k2[i2]
                double sr = 0.0, sim = 0.0;
                const int pb = k * (k - 1) / 2;
                for (int k2 = 1; k2 <= k; ++k2) {
                    const double2 A = Ap[pb + k2 - 1], c = c2b[(k2 - 1) * Lw + i2];
                    sr += A.x * c.x - A.y * c.y;
                    sim += A.x * c.y + A.y * c.x;
                }
                double2 &G1 = g1[(k - 1) * Lw + i1];
                G1.x += w * sr;
                G1.y -= w * sim;
                // slot 2 (k2 = k): sum_{k1 >= k} A_{k1,k} c_k1[i1]
                sr = 0.0;
                sim = 0.0;
                for (int k1 = k; k1 <= K; ++k1) {
                    const double2 A = Ap[k1 * (k1 - 1) / 2 + k - 1], c = c1b[(k1 - 1) * Lw + i1];
                    sr += A.x * c.x - A.y * c.y;
                    sim += A.x * c.y + A.y * c.x;
                }
                double2 &G2 = g2[(k - 1) * Lw + i2];
                G2.x += w * sr;
                G2.y -= w * sim;
            }
        } else {
            // phase 2b: slot-3 partials, thread (chunk, k3)
            const int t = tid - 512;
            const int ch = t / 32, k3 = t % 32 + 1;
            if (ch < kBwdChunks && k3 <= K) {
                double sr = 0.0, sim = 0.0;
                for (int p = ch; p < np; p += kBwdChunks) {
                    const int k1 = pk[p] & 0xff, k2 = pk[p] >> 8;
                    if (k2 < k3) continue;
                    const double g = sgA[ktof(k1, k2, k3)];
                    const double2 T = Ts[p];
                    sr += g * T.x;
                    sim -= g * T.y;   // conj(T)
                }
                g3p[ch * 32 + k3 - 1] = make_double2(sr, sim);
            }
        }
        __syncthreads();
    }
    __syncthreads();
    // write the three blocks
    double2 *out = gpart + (size_t)lt * 3 * K * Lw;
    for (int i = tid; i < 3 * K * Lw; i += NT) out[i] = g1[i];
}

// ---------------------------------------------------------------------------
// gamma[k][l, m] = sum over the (l-triple, slot) blocks whose slot has order
// l, in the fixed order of the list (deterministic).  One wave per (k, l),
// lanes over m = 0..2l; list[lstart[l/2] .. lstart[l/2 + 1]) holds
// lt * 3 + slot; coff[l] is the compact offset of order l.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_grad_reduce(const double2 *__restrict__ gpart, int K, int Lw, int n_c,
                                                    const int *__restrict__ coff, const int *__restrict__ lstart,
                                                    const int *__restrict__ list, double2 *__restrict__ grad)
{
    const int k = blockIdx.x, li = blockIdx.y, l = 2 * li, m = threadIdx.x;
    if (m > 2 * l) return;
    const int q0 = lstart[li], q1 = lstart[li + 1];
    double sr = 0.0, si = 0.0;
    int q = q0;
    for (; q + 4 <= q1; q += 4) {   // four independent loads in flight
        const double2 v0 = gpart[((size_t)list[q] * K + k) * Lw + m];
        const double2 v1 = gpart[((size_t)list[q + 1] * K + k) * Lw + m];
        const double2 v2 = gpart[((size_t)list[q + 2] * K + k) * Lw + m];
        const double2 v3 = gpart[((size_t)list[q + 3] * K + k) * Lw + m];
        sr += v0.x; si += v0.y;
        sr += v1.x; si += v1.y;
        sr += v2.x; si += v2.y;
        sr += v3.x; si += v3.y;
    }
    for (; q < q1; ++q) {
        const double2 v = gpart[((size_t)list[q] * K + k) * Lw + m];
        sr += v.x;
        si += v.y;
    }
    grad[(size_t)k * n_c + coff[l] + m] = make_double2(sr, si);
}

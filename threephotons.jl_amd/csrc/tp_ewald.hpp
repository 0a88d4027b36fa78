// tp_ewald.hpp -- the curved Ewald sphere (lambda != 0) in the c3 contraction.
//
// calculate_basis (correlations.jl:87-153) takes the Legendre factors of
// P[kt, j] = w_j sphPlm(l1, m1, theta_k1) sphPlm(l2, m2, theta_k2)
// sphPlm(l3, mu3, theta_k3) at the shell angles theta_k = acos(k lambda dq /
// (4 pi)) (:94, :142).  At lambda = 0 every theta_k is pi/2, P does not depend
// on the k-triple and G = H . As is one GEMM over l-triples.  Otherwise the
// (class, l-triple) weight depends on the k-triple through three per-shell
// factors, so G is formed column by column:
//   G[c][kt] = sum over entries e of class c of
//              hw_e Y[k1][l1 m1] Y[k2][l2 m2] Y[k3][l3 mu3] As[lt_e][kt],
// with c the folded class {(m2, mu3), (-m2, -mu3)} (both members give the
// same term for even l1 + l2 + l3, so hw_e = 2 w_e, or w_e for (0, 0)).
// The triple invariants As and the contraction c3 = Bm . G are unchanged.
#pragma once

constexpr int kLamMaxThreads = 1024;

// One term of an output row (forward: a class of G; backward: an l-triple of
// dE/dAs): its folded 3j weight, the input row (forward: the l-triple, a row
// of As; backward: the class, a row of dE/dG), the positions r1, r2, r3 of
// its three sphPlm rows (k1, k2 and k3 factors) among the rows staged for
// the output row, and two flags: 1 when r3 differs from the previous
// entry's, 2 when (r2, r3) does (the first entry of a row carries both).
struct LamRec {
    double hw;
    uint32_t a;   // input row | flags << 16
    uint32_t b;   // r1 | r2 << 8 | r3 << 16
};

// Dynamic LDS of k_glam: an output row's staged sphPlm rows and its entries.
inline size_t glam_lds_bytes(int max_rows, int K, int max_ent)
{
    return sizeof(double) * (size_t)((max_rows * K + 1) & ~1) + sizeof(LamRec) * (size_t)max_ent;
}

// One workgroup per (block of blockDim.x columns, output row o): thread ->
// column (columns past klength, and in a batch the columns past a block's
// nvalid, get zeros).
//   Out[o][kt] = sum over entries e of o: hw_e Y[k1][r1_e] Y[k2][r2_e] Y[k3][r3_e] In[in_e][kt].
// Forward (G = H_kt As): o is a class, which fixes m1 = -(m2 + mu3), m2 and
// mu3, so its sphPlm rows are (l, m1), (l, m2), (l, mu3) for l = 0, lstep,
// ..., L.  Backward (dE/dAs = H_kt^T dE/dG): o is an l-triple, whose rows are
// (l1, m), (l2, m), (l3, mu) for every m.  The rows (yrows[ystart[o]] ..,
// indices y = l (2L+1) + m + L of Y[y K + k - 1] = sphPlm(l, m, theta_k))
// and the row's entries are staged in LDS, so the only global load per
// entry is the In value.  ktk[kt] = k1 | k2 << 10 | k3 << 20.  With the
// blocks in x and the rows in y, workgroup i runs on XCD i % 8, so at 4
// column blocks every XCD reads one block of In (3.4 MB of As at L = 18,
// K = 26: it stays in that XCD's L2).
// BATCH: the columns of a batch (As in FP32, as k_triple_sym_batch writes
// it); 64-column block j holds the k-triples bkt0[j] .. bkt0[j] + nvalid[j]
// - 1 of one evaluation (a run's k-triples are the first ones of the basis's
// order, k1 >= k2 >= k3 with k1 outermost).  A batch has thousands of column
// blocks, so its grid has the classes in x and the column blocks in y: the
// workgroups running at one time share one column block of As, which every
// XCD then holds in its L2, and As is read from HBM about once.
template <bool BATCH, typename AT>
__global__ __launch_bounds__(kLamMaxThreads) void k_glam_t(const AT *__restrict__ As, int64_t lda,
                                                           const int *__restrict__ ktk, int klength,
                                                           const double *__restrict__ Y, int K,
                                                           const int *__restrict__ yrows, const int *__restrict__ ystart,
                                                           int max_rows, const LamRec *__restrict__ ent,
                                                           const int *__restrict__ cstart, double *__restrict__ G,
                                                           int64_t ldg, const int *__restrict__ nvalid,
                                                           const int *__restrict__ bkt0)
{
    extern __shared__ double lam_sm[];
    const int c = BATCH ? blockIdx.x : blockIdx.y, cblk = BATCH ? blockIdx.y : blockIdx.x, nt = blockDim.x;
    const int eb = cstart[c], n = cstart[c + 1] - eb;
    const int y0 = ystart[c], ny = ystart[c + 1] - y0;
    double *Ys = lam_sm;   // [ny][K]
    LamRec *se = reinterpret_cast<LamRec *>(lam_sm + ((max_rows * K + 1) & ~1));
    for (int i = threadIdx.x; i < ny * K; i += nt) Ys[i] = Y[(size_t)yrows[y0 + i / K] * K + i % K];
    for (int i = threadIdx.x; i < n; i += nt) se[i] = ent[eb + i];
    __syncthreads();
    const int col = cblk * nt + threadIdx.x;
    if (col >= ldg) return;   // no barriers below
    int kt = col;
    if constexpr (BATCH) {
        const int loc = col & 63;
        kt = loc < nvalid[col >> 6] ? bkt0[col >> 6] + loc : -1;
    } else if (kt >= klength) {
        kt = -1;
    }
    if (kt < 0) {
        G[(int64_t)c * ldg + col] = 0.0;
        return;
    }
    const int p = ktk[kt];
    const double *Y1 = Ys + ((p & 1023) - 1);
    const double *Y2 = Ys + (((p >> 10) & 1023) - 1);
    const double *Y3 = Ys + (((p >> 20) & 1023) - 1);
    const AT *Ak = As + col;
    // entries run sorted by (l3, l2, l1): the k3 factor and the (k2, k3)
    // product are reloaded only when l3 or (l2, l3) changes; two
    // accumulators keep two independent chains of loads and FMAs in flight
    double acc0 = 0.0, acc1 = 0.0, y3 = 0.0, y23 = 0.0;
    int e = 0;
    for (; e + 1 < n; e += 2) {
        const LamRec a = se[e], b = se[e + 1];   // the same for every lane
        if (a.a & 0x10000u) y3 = Y3[(int)(a.b >> 16) * K];
        if (a.a & 0x20000u) y23 = Y2[(int)((a.b >> 8) & 255u) * K] * y3;
        const double ta = a.hw * Y1[(int)(a.b & 255u) * K] * y23;
        if (b.a & 0x10000u) y3 = Y3[(int)(b.b >> 16) * K];
        if (b.a & 0x20000u) y23 = Y2[(int)((b.b >> 8) & 255u) * K] * y3;
        const double tb = b.hw * Y1[(int)(b.b & 255u) * K] * y23;
        acc0 += ta * (double)Ak[(int64_t)(a.a & 0xffffu) * lda];
        acc1 += tb * (double)Ak[(int64_t)(b.a & 0xffffu) * lda];
    }
    if (e < n) {
        const LamRec a = se[e];
        if (a.a & 0x10000u) y3 = Y3[(int)(a.b >> 16) * K];
        if (a.a & 0x20000u) y23 = Y2[(int)((a.b >> 8) & 255u) * K] * y3;
        acc0 += a.hw * Y1[(int)(a.b & 255u) * K] * y23 * (double)Ak[(int64_t)(a.a & 0xffffu) * lda];
    }
    G[(int64_t)c * ldg + col] = acc0 + acc1;
}

#define k_glam (k_glam_t<false, double>)
#define k_glam_batch (k_glam_t<true, float>)

// out[i] = sum_b in[b blk_stride + i], i < n, blocks added in order: the njc
// partial row blocks of dE/dG that the backward contraction leaves.
__global__ __launch_bounds__(256) void k_sum_blocks(const double *__restrict__ in, int nblk, int64_t blk_stride,
                                                    int64_t n, double *__restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += in[b * blk_stride + i];
        out[i] = s;
    }
}

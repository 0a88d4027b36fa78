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

constexpr int kLamThreads = 256;
constexpr int kLamClasses = 1;    // classes per workgroup (grid y)

struct LamEnt {
    double hw;        // folded 3j weight
    int lt;           // l-triple (row of As)
    int y1, y2, y3;   // (l, m) rows of the sphPlm table, times K
    int fl;           // 1: l3 differs from the previous entry's, 2: (l2, l3) differs
};

// One thread per column kt of [0, ldg) (columns past klength get zeros), the
// classes [kLamClasses blockIdx.y, + kLamClasses) of G.  ktk[kt] = k1 | k2 << 10
// | k3 << 20; Y[y + k - 1] = sphPlm(l, m, theta_k) for the (l, m) row y: the
// shells of one (l, m) are contiguous, so the lanes of a wave (consecutive
// k-triples, mostly differing in k3) read one or two cache lines.
__global__ __launch_bounds__(kLamThreads) void k_glam(const double *__restrict__ As, int64_t lda,
                                                      const int *__restrict__ ktk, int klength,
                                                      const double *__restrict__ Y,
                                                      const LamEnt *__restrict__ ent, const int *__restrict__ cstart,
                                                      int n_m, double *__restrict__ G, int64_t ldg)
{
    const int kt = blockIdx.x * kLamThreads + threadIdx.x;
    if (kt >= ldg) return;   // no barriers below
    const int c0 = blockIdx.y * kLamClasses, c1 = min(c0 + kLamClasses, n_m);
    if (kt >= klength) {
        for (int c = c0; c < c1; ++c) G[(int64_t)c * ldg + kt] = 0.0;
        return;
    }
    const int p = ktk[kt];
    const double *Y1 = Y + ((p & 1023) - 1);
    const double *Y2 = Y + (((p >> 10) & 1023) - 1);
    const double *Y3 = Y + (((p >> 20) & 1023) - 1);
    for (int c = c0; c < c1; ++c) {
        // a class's entries run sorted by (l3, l2, l1): the k3 factor and the
        // (k2, k3) product are reloaded only when l3 or (l2, l3) changes (the
        // first entry carries both flags); two accumulators keep two
        // independent chains of loads and FMAs in flight
        double acc0 = 0.0, acc1 = 0.0, y3 = 0.0, y23 = 0.0;
        const int e1 = cstart[c + 1];
        int e = cstart[c];
        for (; e + 1 < e1; e += 2) {
            const LamEnt a = ent[e], b = ent[e + 1];   // the same for every lane
            if (a.fl & 1) y3 = Y3[a.y3];
            if (a.fl & 2) y23 = Y2[a.y2] * y3;
            const double ta = a.hw * Y1[a.y1] * y23;
            if (b.fl & 1) y3 = Y3[b.y3];
            if (b.fl & 2) y23 = Y2[b.y2] * y3;
            const double tb = b.hw * Y1[b.y1] * y23;
            acc0 += ta * As[(int64_t)a.lt * lda + kt];
            acc1 += tb * As[(int64_t)b.lt * lda + kt];
        }
        if (e < e1) {
            const LamEnt a = ent[e];
            if (a.fl & 1) y3 = Y3[a.y3];
            if (a.fl & 2) y23 = Y2[a.y2] * y3;
            acc0 += a.hw * Y1[a.y1] * y23 * As[(int64_t)a.lt * lda + kt];
        }
        G[(int64_t)c * ldg + kt] = acc0 + acc1;
    }
}

/*
 * tp_oracle.c -- CPU FP64 restatement of the ThreePhotons.jl three-photon
 * correlation (c3) and its Bayes log-likelihood energy.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is linked by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * never by the product library (libtp.so).  Every function names the
 * reference file:line it restates (paths relative to the reference root).
 *
 * Pins (what checks each function; the tests live in tests/):
 *   or_wigner3j     exact sympy values (tests/golden/wigner3j.json), closed
 *                   forms, orthogonality sums.
 *   or_sphPlm       mpmath values (tests/golden/sphplm.json), closed form at
 *                   theta=pi/2, addition theorem.
 *   or_seanindex    closed form bijection onto 1..LMAX^2.
 *   basis           enumeration counts (basislen/n_lt at L=8 and L=18).
 *   c3 (3 forms)    dense B*PA == factored F*As == slow flab/c3_slice;
 *                   brute-force SO(3) orientational average of I(q1)I(q2)I(q3);
 *                   SO(3) invariance under Wigner-D; single-shell closed form.
 *   energy          Gibbs equality case, hand-computed 2-bin example, clamp.
 *   moves           two-frame rotation == C^T G C of the QR Haar matrix;
 *                   orthogonality, det, trace, c2 invariance.
 *   philox          Random123 known-answer vectors.
 *   wigner D        orthogonality integral of tests/test_spherical_harmonics.jl,
 *                   group property.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Index helpers                                                             */
/* ------------------------------------------------------------------------ */

/* seanindex: spherical_harmonics.jl:163-170 (1-based position of (l,m) in a
 * length-LMAX^2 coefficient vector; m>=0 block first, then m<0 block). */
int64_t or_seanindex(int l, int m, int LMAX)
{
    int64_t bigL = LMAX - 1;
    if (m >= 0)
        return (int64_t)m * (bigL + 1) - ((int64_t)m * (m - 1)) / 2 + (l - m) + 1;
    return (bigL * (bigL + 3)) / 2 + 1 + ((bigL + m) * (bigL + m + 1)) / 2 + (l + m) + 1;
}

/* tripletFactor: datagen.jl:166-174 */
double or_tripletFactor(int k1, int k2, int k3)
{
    if (k1 == k2 && k2 == k3) return 1.0;
    if ((k1 == k2 && k2 != k3) || (k1 == k3 && k3 != k2) || (k2 == k3 && k3 != k1)) return 3.0;
    return 6.0;
}

/* ------------------------------------------------------------------------ */
/* Wigner 3j: the reference calls GSL gsl_sf_coupling_3j with doubled        */
/* arguments (spherical_harmonics.jl:63-65).  Restated from the Racah        */
/* formula written as an alternating sum of binomial products:               */
/*   3j = (-1)^{j1-j2-m3} S sqrt( prod (j+-m)! / ((J+1)! t1! t2! t3!) ),      */
/*   S  = sum_k (-1)^k C(t1,k) C(t2,j1-m1-k) C(t3,j2+m2-k),                  */
/* t1=j1+j2-j3, t2=j1-j2+j3, t3=-j1+j2+j3.  S is summed exactly in 128-bit   */
/* integers (each binomial is <= C(44,22) < 2^41 for l <= 22, so the triple */
/* products stay below 2^123); the prefactor in long double.                */
/* ------------------------------------------------------------------------ */

static long double or_fact_ld[256];
static int or_fact_ready = 0;

static void or_fact_init(void)
{
    if (or_fact_ready) return;
    or_fact_ld[0] = 1.0L;
    for (int i = 1; i < 256; ++i) or_fact_ld[i] = or_fact_ld[i - 1] * (long double)i;
    or_fact_ready = 1;
}

static __int128 or_binom(int n, int k)
{
    if (k < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    __int128 c = 1;
    for (int i = 1; i <= k; ++i) c = c * (n - k + i) / i;
    return c;
}

double or_wigner3j(int j1, int m1, int j2, int m2, int j3, int m3)
{
    or_fact_init();
    if (m1 + m2 + m3 != 0) return 0.0;
    if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
    if (j3 < abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    if (j1 > 22 || j2 > 22 || j3 > 22) return NAN;   /* outside the exact-sum range */
    int t1 = j1 + j2 - j3, t2 = j1 - j2 + j3, t3 = -j1 + j2 + j3, J = j1 + j2 + j3;
    __int128 S = 0;
    for (int k = 0; k <= t1; ++k) {
        int a = j1 - m1 - k, b = j2 + m2 - k;
        if (a < 0 || a > t2 || b < 0 || b > t3) continue;
        __int128 term = or_binom(t1, k) * or_binom(t2, a) * or_binom(t3, b);
        S += (k & 1) ? -term : term;
    }
    if (S == 0) return 0.0;
    long double num = or_fact_ld[j1 + m1] * or_fact_ld[j1 - m1] * or_fact_ld[j2 + m2] *
                      or_fact_ld[j2 - m2] * or_fact_ld[j3 + m3] * or_fact_ld[j3 - m3];
    long double den = or_fact_ld[J + 1] * or_fact_ld[t1] * or_fact_ld[t2] * or_fact_ld[t3];
    long double v = (long double)S * sqrtl(num / den);
    int ph = j1 - j2 - m3;
    if (ph & 1) v = -v;
    return (double)v;
}

/* ------------------------------------------------------------------------ */
/* sphPlm: spherical_harmonics.jl:69-76 -> GSL gsl_sf_legendre_sphPlm(l,|m|, */
/* cos theta) = sqrt((2l+1)/4pi (l-m)!/(l+m)!) P_l^m(x) with the Condon-     */
/* Shortley phase; for m<0 the reference multiplies by (-1)^|m|.  Restated   */
/* with the standard normalised three-term recurrence in long double.  As in */
/* the reference, x = cos(theta) is formed in double first.                  */
/* ------------------------------------------------------------------------ */
double or_sphPlm(int l, int m, double theta)
{
    int ma = m < 0 ? -m : m;
    if (ma > l) return 0.0;
    long double x = (long double)cos(theta);
    long double s = sqrtl((1.0L - x) * (1.0L + x));
    long double pmm = sqrtl(1.0L / (4.0L * (long double)OR_PI));
    for (int i = 1; i <= ma; ++i) pmm *= -sqrtl((2.0L * i + 1.0L) / (2.0L * i)) * s;
    long double res;
    if (l == ma) {
        res = pmm;
    } else {
        long double p1 = x * sqrtl(2.0L * ma + 3.0L) * pmm;
        if (l == ma + 1) {
            res = p1;
        } else {
            long double p0 = pmm, pl = 0;
            for (int ll = ma + 2; ll <= l; ++ll) {
                long double a = sqrtl((4.0L * ll * ll - 1.0L) / ((long double)ll * ll - (long double)ma * ma));
                long double b = sqrtl((((long double)(ll - 1) * (ll - 1)) - (long double)ma * ma) /
                                      (4.0L * (ll - 1) * (ll - 1) - 1.0L));
                pl = a * (x * p1 - b * p0);
                p0 = p1;
                p1 = pl;
            }
            res = pl;
        }
    }
    if (m < 0 && (ma & 1)) res = -res;
    return (double)res;
}

/* ------------------------------------------------------------------------ */
/* Basis: BasisType / calculate_basis / calculate_combolist                  */
/* (correlations.jl:35-63, 68-84, 87-153).                                   */
/* Storage differs from the reference only where the reference stores data  */
/* that is a pure function of a few indices: B[r,j] = cos(m2 a + mu3 b) is   */
/* kept as unique (m2,mu3) columns plus a per-entry column index, and        */
/* P[kt,j] = w Y Y Y is evaluated from a per-shell Y table.  The arithmetic  */
/* of the dense reference contraction is unchanged.                          */
/* ------------------------------------------------------------------------ */
typedef struct or_basis {
    int N, L, LMAX, K, for_intensity, fp32_tables;
    double lambda, dq;
    int64_t basislen, n_lt, klength, R;
    double *w;          /* [basislen] wignerlist */
    int64_t *ind;       /* [9*basislen] indices (reference layout, 1-based seanindex) */
    int64_t *lt;        /* [5*n_lt] mcombolist rows: l1,l2,l3,mcombos,jstart (jstart 1-based) */
    int64_t *lt_of_j;   /* [basislen] l-triple of each entry (0-based) */
    int n_m;            /* unique (m2,mu3) pairs */
    int *m_of_col;      /* [2*n_m] */
    int *col_of_j;      /* [basislen] */
    double *Bm;         /* [R*n_m] column-major cos(m2 a + mu3 b) */
    int Ltab;           /* max l in Y table */
    double *Y;          /* [K][(Ltab+1)][2*Ltab+1] sphPlm(l,m,theta_k) */
    int64_t *kt_k;      /* [3*klength] k1,k2,k3 (1-based) */
} or_basis;

static double or_round_tab(const or_basis *b, double x)
{
    return b->fp32_tables ? (double)(float)x : x;
}

/* alpharange / alpharange_2pi: correlations.jl:247-256 */
double or_alpha(int i, int N) { return (i + 0.5) * OR_PI / N; }     /* i 0-based, N points on (0,pi) */
double or_beta(int j, int N) { return (j + 0.5) * OR_PI / N; }      /* j 0-based, 2N points on (0,2pi) */

static inline double or_Yk(const or_basis *b, int k /*1-based*/, int l, int m)
{
    int W = 2 * b->Ltab + 1;
    return b->Y[((int64_t)(k - 1) * (b->Ltab + 1) + l) * W + (m + b->Ltab)];
}

void or_basis_free(or_basis *b)
{
    if (!b) return;
    free(b->w); free(b->ind); free(b->lt); free(b->lt_of_j); free(b->m_of_col);
    free(b->col_of_j); free(b->Bm); free(b->Y); free(b->kt_k);
    free(b);
}

or_basis *or_basis_create(int L, int LMAX, int N, int K, double lambda, double dq,
                          int for_intensity, int fp32_tables)
{
    if (L < 0 || N < 1 || K < 1 || LMAX < L + 1) return NULL;
    or_basis *b = (or_basis *)calloc(1, sizeof(or_basis));
    b->N = N; b->L = L; b->LMAX = LMAX; b->K = K; b->lambda = lambda; b->dq = dq;
    b->for_intensity = for_intensity; b->fp32_tables = fp32_tables;
    b->R = 2LL * N * N;
    b->klength = (int64_t)K * (K + 1) * (K + 2) / 6;
    int step = for_intensity ? 2 : 1;
    const double thr = 100.0 * DBL_EPSILON;   /* correlations.jl:92,117: |w| > 100*eps() */
    /* pass 1: count (correlations.jl:92) */
    int64_t cnt = 0, nlt = 0;
    for (int l1 = 0; l1 <= L; l1 += step)
        for (int l2 = 0; l2 <= L; l2 += step)
            for (int l3 = 0; l3 <= L; l3 += step) {
                int64_t c0 = cnt;
                for (int m1 = -l1; m1 <= l1; ++m1)
                    for (int m2 = -l2; m2 <= l2; ++m2) {
                        int m3 = m1 + m2;   /* 3j(l1 m1 l2 m2 l3 -m3) needs m1+m2-m3 = 0 */
                        if (m3 < -l3 || m3 > l3) continue;
                        if (fabs(or_wigner3j(l1, m1, l2, m2, l3, -m3)) > thr) ++cnt;
                    }
                if (cnt > c0) ++nlt;
            }
    b->basislen = cnt; b->n_lt = nlt;
    b->w = (double *)malloc(sizeof(double) * cnt);
    b->ind = (int64_t *)malloc(sizeof(int64_t) * 9 * cnt);
    b->lt = (int64_t *)malloc(sizeof(int64_t) * 5 * nlt);
    b->lt_of_j = (int64_t *)malloc(sizeof(int64_t) * cnt);
    /* pass 2: fill (correlations.jl:107-129); loop order l1,l2,l3,m1,m2,m3 */
    int64_t i = 0, t = 0;
    for (int l1 = 0; l1 <= L; l1 += step)
        for (int l2 = 0; l2 <= L; l2 += step)
            for (int l3 = 0; l3 <= L; l3 += step) {
                int64_t start = i;
                for (int m1 = -l1; m1 <= l1; ++m1)
                    for (int m2 = -l2; m2 <= l2; ++m2) {
                        int m3 = m1 + m2;
                        if (m3 < -l3 || m3 > l3) continue;
                        double w = or_wigner3j(l1, m1, l2, m2, l3, -m3);
                        if (fabs(w) > thr) {
                            b->w[i] = w;
                            int64_t *d = b->ind + 9 * i;
                            d[0] = or_seanindex(l1, m1, LMAX);
                            d[1] = or_seanindex(l2, m2, LMAX);
                            d[2] = or_seanindex(l3, -m3, LMAX);
                            d[3] = l1; d[4] = m1; d[5] = l2; d[6] = m2; d[7] = l3; d[8] = -m3;
                            b->lt_of_j[i] = t;
                            ++i;
                        }
                    }
                if (i > start) {
                    int64_t *r = b->lt + 5 * t;
                    r[0] = l1; r[1] = l2; r[2] = l3; r[3] = i - start; r[4] = start + 1;
                    ++t;
                }
            }
    /* unique (m2, mu3) columns of B (correlations.jl:141 depends only on them) */
    int MW = 2 * L + 1;
    int *map = (int *)malloc(sizeof(int) * MW * MW);
    for (int q = 0; q < MW * MW; ++q) map[q] = -1;
    b->col_of_j = (int *)malloc(sizeof(int) * cnt);
    b->m_of_col = (int *)malloc(sizeof(int) * 2 * MW * MW);
    int nm = 0;
    for (int64_t j = 0; j < cnt; ++j) {
        int m2 = (int)b->ind[9 * j + 6], mu3 = (int)b->ind[9 * j + 8];
        int key = (m2 + L) * MW + (mu3 + L);
        if (map[key] < 0) {
            map[key] = nm;
            b->m_of_col[2 * nm] = m2; b->m_of_col[2 * nm + 1] = mu3;
            ++nm;
        }
        b->col_of_j[j] = map[key];
    }
    free(map);
    b->n_m = nm;
    b->Bm = (double *)malloc(sizeof(double) * b->R * nm);
    for (int c = 0; c < nm; ++c) {
        int m2 = b->m_of_col[2 * c], mu3 = b->m_of_col[2 * c + 1];
        for (int jb = 0; jb < 2 * N; ++jb)
            for (int ia = 0; ia < N; ++ia) {
                /* row r = a + (b-1)*N (Julia reshape of an N x 2N comprehension) */
                double v = cos(m2 * or_alpha(ia, N) + mu3 * or_beta(jb, N));
                b->Bm[(int64_t)c * b->R + ia + (int64_t)jb * N] = or_round_tab(b, v);
            }
    }
    /* Y table: qlist = acos(k*lambda*dq/(4pi)) (correlations.jl:94) */
    b->Ltab = L;
    int W = 2 * L + 1;
    b->Y = (double *)malloc(sizeof(double) * (int64_t)K * (L + 1) * W);
    for (int k = 1; k <= K; ++k) {
        double th = acos(k * lambda * dq / (4.0 * OR_PI));
        for (int l = 0; l <= L; ++l)
            for (int m = -L; m <= L; ++m)
                b->Y[((int64_t)(k - 1) * (L + 1) + l) * W + (m + L)] = (abs(m) <= l) ? or_sphPlm(l, m, th) : 0.0;
    }
    /* k-triple order: for k1=1:K, k2=1:k1, k3=1:k2 (correlations.jl:72-82,142) */
    b->kt_k = (int64_t *)malloc(sizeof(int64_t) * 3 * b->klength);
    int64_t kt = 0;
    for (int k1 = 1; k1 <= K; ++k1)
        for (int k2 = 1; k2 <= k1; ++k2)
            for (int k3 = 1; k3 <= k2; ++k3) {
                b->kt_k[3 * kt] = k1; b->kt_k[3 * kt + 1] = k2; b->kt_k[3 * kt + 2] = k3;
                ++kt;
            }
    return b;
}

void or_basis_info(const or_basis *b, int64_t *out /* basislen, n_lt, klength, R, n_m */)
{
    out[0] = b->basislen; out[1] = b->n_lt; out[2] = b->klength; out[3] = b->R; out[4] = b->n_m;
}

void or_basis_tables(const or_basis *b, double *w, int64_t *ind, int64_t *lt)
{
    if (w) memcpy(w, b->w, sizeof(double) * b->basislen);
    if (ind) memcpy(ind, b->ind, sizeof(int64_t) * 9 * b->basislen);
    if (lt) memcpy(lt, b->lt, sizeof(int64_t) * 5 * b->n_lt);
}

/* P[kt,j] = w_j * sphPlm(l1,m1,th_k1) sphPlm(l2,m2,th_k2) sphPlm(l3,mu3,th_k3) (correlations.jl:142) */
static inline double or_P(const or_basis *b, int64_t kt, int64_t j)
{
    const int64_t *d = b->ind + 9 * j;
    const int64_t *k = b->kt_k + 3 * kt;
    double v = b->w[j] * or_Yk(b, (int)k[0], (int)d[3], (int)d[4]) *
               or_Yk(b, (int)k[1], (int)d[5], (int)d[6]) * or_Yk(b, (int)k[2], (int)d[7], (int)d[8]);
    return or_round_tab(b, v);
}

/* PAcombos (calculate_combolist, correlations.jl:68-84): 9 x (klength*n_lt) */
void or_pacombos(const or_basis *b, int64_t *out)
{
    int64_t n = 0;
    for (int64_t kt = 0; kt < b->klength; ++kt)
        for (int64_t t = 0; t < b->n_lt; ++t) {
            int64_t *o = out + 9 * n;
            o[0] = b->kt_k[3 * kt]; o[1] = b->kt_k[3 * kt + 1]; o[2] = b->kt_k[3 * kt + 2];
            o[3] = kt + 1;
            o[4] = b->lt[5 * t]; o[5] = b->lt[5 * t + 1]; o[6] = b->lt[5 * t + 2];
            o[7] = b->lt[5 * t + 4]; o[8] = b->lt[5 * t + 3];
            ++n;
        }
}

/* ------------------------------------------------------------------------ */
/* Triple products (calculate_triple_products_fast, correlations.jl:157-176) */
/* coeff: complex interleaved, shell k (1-based) at coeff + 2*(k-1)*ld,      */
/* entry seanindex s (1-based) at +2*(s-1).                                  */
/* As(kt,lt) = sum_{j in lt} w_j Re(c_k1[i1] c_k2[i2] c_k3[i3])              */
/* ------------------------------------------------------------------------ */
static double or_As_one(const or_basis *b, const double *coeff, int64_t ld, int64_t kt, int64_t t)
{
    const int64_t *k = b->kt_k + 3 * kt;
    const double *c1 = coeff + 2 * (k[0] - 1) * ld;
    const double *c2 = coeff + 2 * (k[1] - 1) * ld;
    const double *c3 = coeff + 2 * (k[2] - 1) * ld;
    int64_t j0 = b->lt[5 * t + 4] - 1, mc = b->lt[5 * t + 3];
    double As = 0.0;
    for (int64_t n = 0; n < mc; ++n) {
        int64_t j = j0 + n;
        const int64_t *d = b->ind + 9 * j;
        double ar = c1[2 * (d[0] - 1)], ai = c1[2 * (d[0] - 1) + 1];
        double br = c2[2 * (d[1] - 1)], bi = c2[2 * (d[1] - 1) + 1];
        double cr = c3[2 * (d[2] - 1)], ci = c3[2 * (d[2] - 1) + 1];
        /* Julia: ck1*ck2*ck3 evaluates left to right */
        double pr = ar * br - ai * bi, pi = ar * bi + ai * br;
        As += b->w[j] * (pr * cr - pi * ci);
    }
    return As;
}

/* As table [klength x n_lt] row-major (kt major) */
void or_triple_products(const or_basis *b, const double *coeff, int64_t ld, double *As, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t kt = 0; kt < b->klength; ++kt)
        for (int64_t t = 0; t < b->n_lt; ++t)
            As[kt * b->n_lt + t] = or_As_one(b, coeff, ld, kt, t);
}

/* Dense reference contraction c3raw = B * PA (FullCorrelation_parallized,    */
/* correlations.jl:179-184), PA[j,kt] = As(kt,lt(j)) * P[kt,j] (:173).        */
/* Computes columns kt in [col_lo, col_hi); out is R x (col_hi-col_lo)        */
/* column-major.  When the basis was made with fp32_tables the B and P        */
/* entries are FP32-rounded as the reference stores them (:100-101).          */
void or_c3_dense(const or_basis *b, const double *coeff, int64_t ld, int64_t col_lo, int64_t col_hi,
                 double *out, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
    int64_t R = b->R;
#pragma omp parallel
    {
        double *As = (double *)malloc(sizeof(double) * b->n_lt);
#pragma omp for schedule(dynamic, 1)
        for (int64_t kt = col_lo; kt < col_hi; ++kt) {
            for (int64_t t = 0; t < b->n_lt; ++t) As[t] = or_As_one(b, coeff, ld, kt, t);
            double *o = out + (kt - col_lo) * R;
            for (int64_t r = 0; r < R; ++r) o[r] = 0.0;
            for (int64_t j = 0; j < b->basislen; ++j) {
                double pa = As[b->lt_of_j[j]] * or_P(b, kt, j);
                const double *Bc = b->Bm + (int64_t)b->col_of_j[j] * R;
                for (int64_t r = 0; r < R; ++r) o[r] += Bc[r] * pa;
            }
        }
        free(As);
    }
}

/* Factored contraction (lambda = 0 only): every theta_k = pi/2, so P[kt,j] is */
/* kt-independent and c3raw[:,kt] = sum_lt F[:,lt] As(kt,lt) with             */
/* F[r,lt] = sum_{j in lt} B[r,j] P[j].  This is the grouping the reference    */
/* itself uses in its slow path (flab, correlations.jl:337-362).              */
/* F out: R x n_lt column-major (optional).                                   */
int or_factor_F(const or_basis *b, double *F)
{
    if (b->lambda != 0.0) return -1;
    int64_t R = b->R;
    memset(F, 0, sizeof(double) * R * b->n_lt);
    for (int64_t j = 0; j < b->basislen; ++j) {
        double p = or_P(b, 0, j);
        double *Fc = F + b->lt_of_j[j] * R;
        const double *Bc = b->Bm + (int64_t)b->col_of_j[j] * R;
        for (int64_t r = 0; r < R; ++r) Fc[r] += Bc[r] * p;
    }
    return 0;
}

int or_c3_factored(const or_basis *b, const double *coeff, int64_t ld, double *out, int nthreads)
{
    if (b->lambda != 0.0) return -1;
    int64_t R = b->R, nlt = b->n_lt, KL = b->klength;
    double *F = (double *)malloc(sizeof(double) * R * nlt);
    double *As = (double *)malloc(sizeof(double) * KL * nlt);
    or_factor_F(b, F);
    or_triple_products(b, coeff, ld, As, nthreads);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static)
    for (int64_t kt = 0; kt < KL; ++kt) {
        double *o = out + kt * R;
        for (int64_t r = 0; r < R; ++r) o[r] = 0.0;
        for (int64_t t = 0; t < nlt; ++t) {
            double a = As[kt * nlt + t];
            const double *Fc = F + t * R;
            for (int64_t r = 0; r < R; ++r) o[r] += Fc[r] * a;
        }
    }
    free(F); free(As);
    return 0;
}

/* Slow formulation: flab (correlations.jl:337-362) and c3_slice (:369-391)   */
/* used with the 11-argument form (the 5-argument call at :406 is broken).    */
static double or_flab(int k1, int k2, int k3, double dq, double lambda, int l1, int l2, int l3, double a, double bb)
{
    double th1 = acos(k1 * lambda * dq / (4 * OR_PI));
    double th2 = acos(k2 * lambda * dq / (4 * OR_PI));
    double th3 = acos(k3 * lambda * dq / (4 * OR_PI));
    if (!(abs(l1 - l2) <= l3 && l3 <= l1 + l2)) return 0.0;
    double fr = 0.0;
    for (int m1p = -l1; m1p <= l1; ++m1p)
        for (int m2p = -l2; m2p <= l2; ++m2p)
            for (int m3p = -l3; m3p <= l3; ++m3p) {
                double fac = or_wigner3j(l1, m1p, l2, m2p, l3, -m3p);
                if (fabs(fac) > 100 * DBL_EPSILON) {
                    double y = or_sphPlm(l1, m1p, th1) * or_sphPlm(l2, m2p, th2) * or_sphPlm(l3, -m3p, th3);
                    fr += fac * y * cos(m2p * a - m3p * bb);   /* real part of fac*y*exp(i(m2'a - m3'b)) */
                }
            }
    return fr;
}

/* one slice (N x 2N, column-major) for shells k1>=k2>=k3 (1-based) */
void or_c3_slice(const double *coeff, int64_t ld, int N, int L, int LMAX, int k1, int k2, int k3,
                 double dq, double lambda, double *out)
{
    const double *ck1 = coeff + 2 * (int64_t)(k1 - 1) * ld;
    const double *ck2 = coeff + 2 * (int64_t)(k2 - 1) * ld;
    const double *ck3 = coeff + 2 * (int64_t)(k3 - 1) * ld;
    int R = 2 * N * N;
    double *m = (double *)malloc(sizeof(double) * R);
    memset(out, 0, sizeof(double) * R);
    for (int l1 = 0; l1 <= L; l1 += 2)
        for (int l2 = 0; l2 <= L; l2 += 2)
            for (int l3 = 0; l3 <= L; l3 += 2) {
                double nrm = 0.0;
                for (int jb = 0; jb < 2 * N; ++jb)
                    for (int ia = 0; ia < N; ++ia) {
                        double v = or_flab(k1, k2, k3, dq, lambda, l1, l2, l3, or_alpha(ia, N), or_beta(jb, N));
                        m[ia + jb * N] = v;
                        nrm += v * v;
                    }
                if (!(sqrt(nrm) > 3 * DBL_EPSILON)) continue;
                double sr = 0.0;   /* real part of sum w c c c (the sum is real for real intensities;
                                      the reference takes real(slice) at the end, :390) */
                double si = 0.0;
                for (int m1 = -l1; m1 <= l1; ++m1)
                    for (int m2 = -l2; m2 <= l2; ++m2)
                        for (int m3 = -l3; m3 <= l3; ++m3) {
                            double w = or_wigner3j(l1, m1, l2, m2, l3, -m3);
                            if (fabs(w) > 3 * DBL_EPSILON) {
                                const double *a = ck1 + 2 * (or_seanindex(l1, m1, LMAX) - 1);
                                const double *b2 = ck2 + 2 * (or_seanindex(l2, m2, LMAX) - 1);
                                const double *c = ck3 + 2 * (or_seanindex(l3, -m3, LMAX) - 1);
                                double pr = a[0] * b2[0] - a[1] * b2[1], pi = a[0] * b2[1] + a[1] * b2[0];
                                sr += w * (pr * c[0] - pi * c[1]);
                                si += w * (pr * c[1] + pi * c[0]);
                            }
                        }
                (void)si;
                for (int r = 0; r < R; ++r) out[r] += m[r] * sr;
            }
    free(m);
}

/* ------------------------------------------------------------------------ */
/* Bayes energy: energy(), correlations.jl:201-226.                          */
/* c3raw: R x klength column-major (return_raw layout).                      */
/* c3ref: Julia C3 array N x 2N x Kr x Kr x Kr column-major.                 */
/* The column index i walks K3_range = k3lo:k3hi exactly as the reference    */
/* does (reproducing its running-index behaviour when k3lo > 1).             */
/* ------------------------------------------------------------------------ */
double or_energy(const double *c3raw, int N, int64_t ncols, const double *c3ref, int Kr, int k3lo, int k3hi)
{
    int64_t R = 2LL * N * N;
    double res = 0.0;
    int64_t i = 0;
    for (int k1 = k3lo; k1 <= k3hi; ++k1)
        for (int k2 = k3lo; k2 <= k1; ++k2)
            for (int k3 = k3lo; k3 <= k2; ++k3) {
                if (i >= ncols) return NAN;
                const double *pc = c3raw + i * R;
                ++i;
                const double *q = c3ref + (int64_t)R * ((k3 - 1) + (int64_t)Kr * ((k2 - 1) + (int64_t)Kr * (k1 - 1)));
                double sp = 0.0, sq = 0.0;
                for (int64_t r = 0; r < R; ++r) {
                    double p = pc[r] > 1e-30 ? pc[r] : 1e-30;   /* max.(c3, 1e-30), :203 */
                    sp += fabs(p);
                    sq += fabs(q[r]);
                }
                double acc = 0.0;
                for (int64_t r = 0; r < R; ++r) {
                    double p = pc[r] > 1e-30 ? pc[r] : 1e-30;
                    acc += (q[r] / sq) * log(p / sp);
                }
                res += -1.0 * acc * or_tripletFactor(k1, k2, k3);
            }
    return res;
}

/* ------------------------------------------------------------------------ */
/* Gradient of the Bayes energy with respect to the coefficients (no         */
/* counterpart in the reference, which only evaluates the energy; it is the  */
/* chain rule through energy() (correlations.jl:201-226), the factored        */
/* contraction and the triple products (:165-174)):                           */
/*   D[r,i]  = dE/dc3raw[r,i] = tF (sum_r' qhat_r' / S - qhat_r / p_r)        */
/*             for c3raw[r,i] > 1e-30 (0 where the clamp is active),          */
/*             qhat = q / sum|q|, S = sum_r p_r;                              */
/*   gAs(kt,lt) = sum_r F[r,lt] D[r,kt];                                      */
/*   gamma_k[i] = dE/dRe c_k[i] + i dE/dIm c_k[i] accumulates, for every     */
/*   entry j of lt and kt = (k1,k2,k3), gAs w_j conj(product of the other two */
/*   factors) in each of the three slots (d Re(abc) = Re(conj(conj(bc)) da)). */
/* Every coefficient entry is treated as an independent complex variable.    */
/* grad: K shells x ld complex (interleaved), zeroed here.  Returns E.       */
/* ------------------------------------------------------------------------ */
double or_energy_grad(const or_basis *b, const double *coeff, int64_t ld, const double *c3ref, int Kr, int k3lo,
                      int k3hi, double *grad, int nthreads)
{
    if (b->lambda != 0.0) return NAN;
    const int64_t R = b->R, nlt = b->n_lt, KL = b->klength;
    double *c3 = (double *)malloc(sizeof(double) * R * KL);
    double *F = (double *)malloc(sizeof(double) * R * nlt);
    double *D = (double *)calloc((size_t)R * KL, sizeof(double));
    double *gA = (double *)malloc(sizeof(double) * KL * nlt);
    or_c3_factored(b, coeff, ld, c3, nthreads);
    or_factor_F(b, F);
    const double E = or_energy(c3, b->N, KL, c3ref, Kr, k3lo, k3hi);
    /* D along energy()'s column walk */
    int64_t i = 0;
    for (int k1 = k3lo; k1 <= k3hi; ++k1)
        for (int k2 = k3lo; k2 <= k1; ++k2)
            for (int k3 = k3lo; k3 <= k2; ++k3, ++i) {
                const double *pc = c3 + i * R;
                const double *q = c3ref + (int64_t)R * ((k3 - 1) + (int64_t)Kr * ((k2 - 1) + (int64_t)Kr * (k1 - 1)));
                double sp = 0.0, sq = 0.0, qs = 0.0;
                for (int64_t r = 0; r < R; ++r) {
                    sp += pc[r] > 1e-30 ? pc[r] : 1e-30;
                    sq += fabs(q[r]);
                }
                for (int64_t r = 0; r < R; ++r) qs += q[r] / sq;
                const double tf = or_tripletFactor(k1, k2, k3);
                for (int64_t r = 0; r < R; ++r)
                    D[i * R + r] = pc[r] > 1e-30 ? tf * (qs / sp - (q[r] / sq) / pc[r]) : 0.0;
            }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(static)
    for (int64_t kt = 0; kt < KL; ++kt)
        for (int64_t t = 0; t < nlt; ++t) {
            double s = 0.0;
            for (int64_t r = 0; r < R; ++r) s += F[t * R + r] * D[kt * R + r];
            gA[kt * nlt + t] = s;
        }
    memset(grad, 0, sizeof(double) * 2 * (size_t)b->K * ld);
    for (int64_t kt = 0; kt < KL; ++kt) {
        const int64_t *k = b->kt_k + 3 * kt;
        const double *c1 = coeff + 2 * (k[0] - 1) * ld, *c2 = coeff + 2 * (k[1] - 1) * ld,
                     *c3c = coeff + 2 * (k[2] - 1) * ld;
        double *g1 = grad + 2 * (k[0] - 1) * ld, *g2 = grad + 2 * (k[1] - 1) * ld, *g3 = grad + 2 * (k[2] - 1) * ld;
        for (int64_t t = 0; t < nlt; ++t) {
            const double ga = gA[kt * nlt + t];
            if (ga == 0.0) continue;
            const int64_t j0 = b->lt[5 * t + 4] - 1, mc = b->lt[5 * t + 3];
            for (int64_t n = 0; n < mc; ++n) {
                const int64_t j = j0 + n;
                const int64_t *d = b->ind + 9 * j;
                const int64_t a1 = 2 * (d[0] - 1), a2 = 2 * (d[1] - 1), a3 = 2 * (d[2] - 1);
                const double ar = c1[a1], ai = c1[a1 + 1], br = c2[a2], bi = c2[a2 + 1], cr = c3c[a3], ci = c3c[a3 + 1];
                const double g = ga * b->w[j];
                /* slot 1: conj(b c) */
                g1[a1] += g * (br * cr - bi * ci);
                g1[a1 + 1] -= g * (br * ci + bi * cr);
                /* slot 2: conj(a c) */
                g2[a2] += g * (ar * cr - ai * ci);
                g2[a2 + 1] -= g * (ar * ci + ai * cr);
                /* slot 3: conj(a b) */
                g3[a3] += g * (ar * br - ai * bi);
                g3[a3 + 1] -= g * (ar * bi + ai * br);
            }
        }
    }
    free(c3); free(F); free(D); free(gA);
    return E;
}

/* ------------------------------------------------------------------------ */
/* Photon-pair and photon-triplet histograms of sparse images:               */
/* histogramCorrelationsInPicture_alltoall (datagen.jl:228-282) with         */
/* angle_between (utilities.jl:61-72), doubletFactor (datagen.jl:177-183),   */
/* tripletFactor (:166-174).  Per picture: keep the photons with             */
/* k = round(|p|/dq) in 1..K2 (round = nearest, ties to even, as Julia's     */
/* round), stable-sort them by |p| ascending, then for i > j > k (so         */
/* k1 >= k2 >= k3):                                                          */
/*   c1[k1] += 1/k1;                                                         */
/*   alpha = angle_between(p_i, p_j) in [0, 2pi); alpha_s = 2pi - alpha if   */
/*   alpha > pi; ai = mod(floor(alpha_s/da), N) + 1, da = pi/N;             */
/*   c2[ai,k2,k1] += 1/(doubletFactor k1 k2);                                */
/*   if k1,k2,k3 <= K3: beta = angle_between(p_i, p_k), beta = 2pi - beta if */
/*   alpha > pi; bi = mod(floor(beta/da), 2N) + 1;                           */
/*   c3[ai,bi,k3,k2,k1] += 1/(tripletFactor k1 k2 k3).                       */
/* angle_between returns 0 when the normalised dot product is exactly 0      */
/* (orthogonal photons), as the reference does.  |p| is sqrt(x^2+y^2+z^2)    */
/* (Julia's norm rescales by max|x_i| first; the two agree but in the last   */
/* bit).  Layouts (Julia column-major): c1[K2]; c2[N,K2,K2]; c3[N,2N,K3,K3,K3]; */
/* n1/n2/n3 (may be NULL) count the contributions with the same indices.     */
/* ------------------------------------------------------------------------ */
static double or_doubletFactor(int k1, int k2) { return k1 != k2 ? 2.0 : 1.0; }

double or_angle_between(const double *p1, const double *p2)
{
    const double n1 = sqrt(p1[0] * p1[0] + p1[1] * p1[1] + p1[2] * p1[2]);
    const double n2 = sqrt(p2[0] * p2[0] + p2[1] * p2[1] + p2[2] * p2[2]);
    double val = (p1[0] * p2[0] + p1[1] * p2[1] + p1[2] * p2[2]) / (n1 * n2);
    val = val < -1.0 ? -1.0 : (val > 1.0 ? 1.0 : val);
    if (fabs(val) > 0.0) {
        double angle = acos(val);
        if (p1[0] * p2[1] - p1[1] * p2[0] > 0) angle = 2.0 * OR_PI - angle;
        return angle;
    }
    return 0.0;
}

static int64_t or_fmod_floor(double x, int64_t m)
{
    int64_t f = (int64_t)floor(x);
    int64_t r = f % m;
    return r < 0 ? r + m : r;
}

void or_hist_pictures(const double *photons /* [n][3] */, const int64_t *offsets /* [npic+1] */, int npic,
                      double dq, int N, int K2, int K3, double *c1, double *c2, double *c3, uint64_t *n1,
                      uint64_t *n2, uint64_t *n3)
{
    const double da = OR_PI / N;
    for (int pic = 0; pic < npic; ++pic) {
        const double *ph = photons + 3 * offsets[pic];
        const int64_t np = offsets[pic + 1] - offsets[pic];
        int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (np > 0 ? np : 1));
        double *nrm = (double *)malloc(sizeof(double) * (np > 0 ? np : 1));
        int *kk = (int *)malloc(sizeof(int) * (np > 0 ? np : 1));
        int64_t l = 0;
        for (int64_t i = 0; i < np; ++i) {
            const double *p = ph + 3 * i;
            const double nm = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            const double kr = nearbyint(nm / dq);
            if (kr >= 1.0 && kr <= (double)K2) {
                idx[l] = i;
                nrm[l] = nm;
                kk[l] = (int)kr;
                ++l;
            }
        }
        if (l >= 3) {
            /* stable insertion sort by norm ascending */
            for (int64_t a = 1; a < l; ++a) {
                const int64_t ti = idx[a];
                const double tn = nrm[a];
                const int tk = kk[a];
                int64_t b = a - 1;
                while (b >= 0 && nrm[b] > tn) {
                    idx[b + 1] = idx[b]; nrm[b + 1] = nrm[b]; kk[b + 1] = kk[b];
                    --b;
                }
                idx[b + 1] = ti; nrm[b + 1] = tn; kk[b + 1] = tk;
            }
            for (int64_t i = 0; i < l; ++i) {
                const double *p1 = ph + 3 * idx[i];
                const int k1 = kk[i];
                c1[k1 - 1] += 1.0 / k1;
                if (n1) n1[k1 - 1] += 1;
                for (int64_t j = 0; j < i; ++j) {
                    const double *p2 = ph + 3 * idx[j];
                    const int k2 = kk[j];
                    const double alpha = or_angle_between(p1, p2);
                    const double alpha_s = alpha > OR_PI ? 2.0 * OR_PI - alpha : alpha;
                    const int64_t ai = or_fmod_floor(alpha_s / da, N);
                    const int64_t o2 = ai + (int64_t)N * ((k2 - 1) + (int64_t)K2 * (k1 - 1));
                    c2[o2] += 1.0 / (or_doubletFactor(k1, k2) * k1 * k2);
                    if (n2) n2[o2] += 1;
                    for (int64_t k = 0; k < j; ++k) {
                        const int k3 = kk[k];
                        if (k1 <= K3 && k2 <= K3 && k3 <= K3) {
                            const double *p3 = ph + 3 * idx[k];
                            double beta = or_angle_between(p1, p3);
                            if (alpha > OR_PI) beta = 2.0 * OR_PI - beta;
                            const int64_t bi = or_fmod_floor(beta / da, 2 * (int64_t)N);
                            const int64_t o3 = ai + (int64_t)N * (bi + 2 * (int64_t)N * ((k3 - 1) +
                                               (int64_t)K3 * ((k2 - 1) + (int64_t)K3 * (k1 - 1))));
                            c3[o3] += 1.0 / (or_tripletFactor(k1, k2, k3) * k1 * k2 * k3);
                            if (n3) n3[o3] += 1;
                        }
                    }
                }
            }
        }
        free(idx); free(nrm); free(kk);
    }
}

/* ------------------------------------------------------------------------ */
/* Umat: spherical_harmonics.jl:445-457.  out: (2l+1)^2 complex, row-major   */
/* [u+l][m+l], interleaved re/im.                                             */
/* ------------------------------------------------------------------------ */
void or_umat(int l, double *out)
{
    int n = 2 * l + 1;
    const double s = 1.0 / sqrt(2.0);
    memset(out, 0, sizeof(double) * 2 * n * n);
    for (int u = -l; u <= l; ++u)
        for (int m = -l; m <= l; ++m) {
            double re = 0, im = 0;
            double sg = (abs(m) & 1) ? -1.0 : 1.0;   /* (-1)^m */
            if (m == 0 && u == 0) re += 1.0;
            if (u > 0 && m == u) re += s;
            if (u < 0 && m == u) im += s * sg;
            if (u < 0 && m == -u) im += -s;
            if (u > 0 && m == -u) re += s * sg;
            out[2 * ((u + l) * n + (m + l))] = re;
            out[2 * ((u + l) * n + (m + l)) + 1] = im;
        }
}

/* ------------------------------------------------------------------------ */
/* Rotations.  random_rotation (utilities.jl:165-176): Q,R = qr(A) of a      */
/* Gaussian matrix, C = Q*diag(sign(diag R)), reject det(C) <= 0.            */
/* Householder QR here (row-major n x n).  Returns det sign of C.            */
/* ------------------------------------------------------------------------ */
int or_haar_from_gaussian(int n, const double *A_in, double *C)
{
    double *A = (double *)malloc(sizeof(double) * n * n);
    double *Q = (double *)malloc(sizeof(double) * n * n);
    memcpy(A, A_in, sizeof(double) * n * n);
    for (int i = 0; i < n * n; ++i) Q[i] = 0;
    for (int i = 0; i < n; ++i) Q[i * n + i] = 1;
    double *v = (double *)malloc(sizeof(double) * n);
    int detsign = 1;
    for (int k = 0; k < n - 1; ++k) {
        double nrm = 0;
        for (int i = k; i < n; ++i) nrm += A[i * n + k] * A[i * n + k];
        nrm = sqrt(nrm);
        double alpha = A[k * n + k] > 0 ? -nrm : nrm;
        for (int i = 0; i < n; ++i) v[i] = 0;
        for (int i = k; i < n; ++i) v[i] = A[i * n + k];
        v[k] -= alpha;
        double vn = 0;
        for (int i = k; i < n; ++i) vn += v[i] * v[i];
        if (vn == 0) continue;
        detsign = -detsign;   /* each Householder reflector has det -1 */
        for (int j = 0; j < n; ++j) {     /* A = H A */
            double d = 0;
            for (int i = k; i < n; ++i) d += v[i] * A[i * n + j];
            d = 2 * d / vn;
            for (int i = k; i < n; ++i) A[i * n + j] -= d * v[i];
        }
        for (int i = 0; i < n; ++i) {     /* Q = Q H */
            double d = 0;
            for (int j = k; j < n; ++j) d += Q[i * n + j] * v[j];
            d = 2 * d / vn;
            for (int j = k; j < n; ++j) Q[i * n + j] -= d * v[j];
        }
    }
    /* C = Q diag(sign(R_ii)) */
    for (int j = 0; j < n; ++j) {
        double sg = A[j * n + j] >= 0 ? 1.0 : -1.0;
        if (sg < 0) detsign = -detsign;
        for (int i = 0; i < n; ++i) C[i * n + j] = Q[i * n + j] * sg;
    }
    free(A); free(Q); free(v);
    return detsign;
}

/* random_rotation_step (utilities.jl:180-190): B = inv(C) G(beta) C with     */
/* G the planar rotation in coordinates 1,2.  C orthogonal so inv(C) = C^T.   */
void or_rotation_step_matrix(int n, const double *C, double beta, double *Bout)
{
    double cb = cos(beta), sb = sin(beta);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            /* (C^T G C)_ij = sum_ab C_ai G_ab C_bj */
            double s = 0;
            for (int a = 0; a < n; ++a) {
                double ga0 = (a == 0) ? cb : (a == 1 ? sb : 0.0);    /* G[a][0] */
                double ga1 = (a == 0) ? -sb : (a == 1 ? cb : 0.0);   /* G[a][1] */
                double cai = C[a * n + i];
                if (a >= 2) { s += cai * C[a * n + j]; continue; }
                s += cai * (ga0 * C[0 * n + j] + ga1 * C[1 * n + j]);
            }
            Bout[i * n + j] = s;
        }
}

/* Two-frame form of the same move: with u, v the first two rows of C,       */
/* C^T G C = I + (cos b - 1)(u u^T + v v^T) + sin b (v u^T - u v^T).          */
void or_frame_rotation_matrix(int n, const double *u, const double *v, double beta, double *Bout)
{
    double cm = cos(beta) - 1.0, sb = sin(beta);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            Bout[i * n + j] = (i == j ? 1.0 : 0.0) + cm * (u[i] * u[j] + v[i] * v[j]) + sb * (v[i] * u[j] - u[i] * v[j]);
}

/* Apply a real rotation R (n x n row-major) of order l to the complex        */
/* coefficient vectors of shells 1..nshell: x <- rtc R ctr x                  */
/* (get_MC_step, determination.jl:232-241; cvec_get/set                       */
/* spherical_harmonics.jl:187-199).                                           */
void or_apply_real_rotation(int l, const double *Rm, double *coeff, int64_t ld, int LMAX, int nshell)
{
    int n = 2 * l + 1;
    double *U = (double *)malloc(sizeof(double) * 2 * n * n);
    or_umat(l, U);
    double *x = (double *)malloc(sizeof(double) * 2 * n), *y = (double *)malloc(sizeof(double) * 2 * n),
           *z = (double *)malloc(sizeof(double) * 2 * n);
    for (int k = 0; k < nshell; ++k) {
        double *ck = coeff + 2 * (int64_t)k * ld;
        for (int m = -l; m <= l; ++m) {
            int64_t s = or_seanindex(l, m, LMAX) - 1;
            x[2 * (m + l)] = ck[2 * s]; x[2 * (m + l) + 1] = ck[2 * s + 1];
        }
        /* y = ctr x */
        for (int u = 0; u < n; ++u) {
            double re = 0, im = 0;
            for (int m = 0; m < n; ++m) {
                double ur = U[2 * (u * n + m)], ui = U[2 * (u * n + m) + 1];
                re += ur * x[2 * m] - ui * x[2 * m + 1];
                im += ur * x[2 * m + 1] + ui * x[2 * m];
            }
            y[2 * u] = re; y[2 * u + 1] = im;
        }
        /* z = R y */
        for (int i = 0; i < n; ++i) {
            double re = 0, im = 0;
            for (int j = 0; j < n; ++j) { re += Rm[i * n + j] * y[2 * j]; im += Rm[i * n + j] * y[2 * j + 1]; }
            z[2 * i] = re; z[2 * i + 1] = im;
        }
        /* x = ctr^H z */
        for (int m = 0; m < n; ++m) {
            double re = 0, im = 0;
            for (int u = 0; u < n; ++u) {
                double ur = U[2 * (u * n + m)], ui = -U[2 * (u * n + m) + 1];
                re += ur * z[2 * u] - ui * z[2 * u + 1];
                im += ur * z[2 * u + 1] + ui * z[2 * u];
            }
            int64_t s = or_seanindex(l, m - l, LMAX) - 1;
            ck[2 * s] = re; ck[2 * s + 1] = im;
        }
    }
    free(U); free(x); free(y); free(z);
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as */
/* 1, 2, 3"), the counter-based generator shared with the device code so     */
/* that CPU and GPU trajectories draw identical numbers.                     */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

#define OR_TAG_MOVE 0x4D4F5645u   /* proposal normals */
#define OR_TAG_ACC  0x41434350u   /* Metropolis uniforms */
#define OR_TAG_TEMP 0x54454D50u   /* temperature-probe proposal normals */

static double or_u01(uint32_t lo, uint32_t hi)
{
    uint64_t u = ((uint64_t)hi << 32) | lo;
    return (double)(u >> 11) * (1.0 / 9007199254740992.0);
}

/* normal number j of stream (seed, a=step, b=l-index, tag) via Box-Muller   */
double or_normal(uint64_t seed, uint32_t step, uint32_t lidx, uint32_t j, uint32_t tag)
{
    uint32_t ctr[4] = {j >> 1, lidx, step, tag}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    or_philox4x32_10(ctr, key, o);
    double u1 = or_u01(o[0], o[1]), u2 = or_u01(o[2], o[3]);
    double r = sqrt(-2.0 * log(1.0 - u1));
    double ang = 2.0 * OR_PI * u2;
    return (j & 1) ? r * sin(ang) : r * cos(ang);
}

double or_uniform(uint64_t seed, uint32_t step, uint32_t tag)
{
    uint32_t ctr[4] = {0, 0xFFFFFFFFu, step, tag}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, o[4];
    or_philox4x32_10(ctr, key, o);
    return or_u01(o[0], o[1]);
}

/* The proposal of step `step`: for every even l in 2..L a planar rotation   */
/* by beta = stepsize_l * z in the plane of a uniformly random orthonormal   */
/* 2-frame (u,v) (Gram-Schmidt of two Gaussian vectors, i.e. the first two   */
/* columns of the sign-fixed QR factor of a Gaussian matrix), applied to all */
/* nshell shells.  Equal in distribution to get_MC_step                      */
/* (determination.jl:232-241), which uses two rows of a Haar SO(2l+1) matrix.*/
void or_mc_propose(uint64_t seed, uint32_t step, uint32_t tag, int L, const double *stepsizes /* [L/2], index l/2-1 */,
                   const double *coeff_in, double *coeff_out, int64_t ld, int LMAX, int nshell)
{
    memcpy(coeff_out, coeff_in, sizeof(double) * 2 * ld * nshell);
    for (int l = 2; l <= L; l += 2) {
        int n = 2 * l + 1, li = l / 2 - 1;
        double *u = (double *)malloc(sizeof(double) * n), *v = (double *)malloc(sizeof(double) * n);
        double *Rm = (double *)malloc(sizeof(double) * n * n);
        for (int i = 0; i < n; ++i) u[i] = or_normal(seed, step, li, i, tag);
        for (int i = 0; i < n; ++i) v[i] = or_normal(seed, step, li, n + i, tag);
        double z = or_normal(seed, step, li, 2 * n, tag);
        double nu = 0; for (int i = 0; i < n; ++i) nu += u[i] * u[i];
        nu = sqrt(nu); for (int i = 0; i < n; ++i) u[i] /= nu;
        double d = 0; for (int i = 0; i < n; ++i) d += u[i] * v[i];
        for (int i = 0; i < n; ++i) v[i] -= d * u[i];
        double nv = 0; for (int i = 0; i < n; ++i) nv += v[i] * v[i];
        nv = sqrt(nv); for (int i = 0; i < n; ++i) v[i] /= nv;
        or_frame_rotation_matrix(n, u, v, stepsizes[li] * z, Rm);
        or_apply_real_rotation(l, Rm, coeff_out, ld, LMAX, nshell);
        free(u); free(v); free(Rm);
    }
}

/* ------------------------------------------------------------------------ */
/* Plain annealing driver: rotate_all_at_once (determination.jl:281-338),    */
/* calculate_temperature (:244-261), evaluate_step (:491-515),               */
/* increase/decrease_stepsize (:478-489).  A fixed number of steps replaces  */
/* the step-size stopping rule, as the benchmark configurations specify.     */
/* Energies use the factored contraction (lambda = 0).                       */
/* ------------------------------------------------------------------------ */
static double or_energy_of(const or_basis *b, const double *coeff, int64_t ld, const double *c3ref, double *c3buf, int nthreads)
{
    or_c3_factored(b, coeff, ld, c3buf, nthreads);
    return or_energy(c3buf, b->N, b->klength, c3ref, b->K, 1, b->K);
}

int or_anneal(const or_basis *b, const double *c3ref, double *coeff /* in/out, KMAX shells */, int64_t ld, int KMAX,
              int nsteps, double initial_stepsize, double stepsizefactor, double initial_temperature_factor,
              double temperature_decay, uint64_t seed, int ntemp,
              double *E_trace /* [nsteps] E after step */, double *Enew_trace /* [nsteps] */, int *acc_trace,
              double *out_scalars /* E0, T0, E_final, T_final */, int nthreads)
{
    if (b->lambda != 0.0) return -1;
    int L = b->L, nl = L / 2;
    double *st = (double *)malloc(sizeof(double) * (nl > 0 ? nl : 1));
    for (int l = 2; l <= L; l += 2) st[l / 2 - 1] = initial_stepsize * (l - 1);   /* :295 */
    double *c3buf = (double *)malloc(sizeof(double) * b->R * b->klength);
    double *step = (double *)malloc(sizeof(double) * 2 * ld * KMAX);
    double E = or_energy_of(b, coeff, ld, c3ref, c3buf, nthreads);
    /* calculate_temperature: std of `ntemp` proposal energies x factor */
    double T = 0.0;
    if (initial_temperature_factor > 0.0 && ntemp > 1) {
        double s1 = 0, s2 = 0;
        double *es = (double *)malloc(sizeof(double) * ntemp);
        for (int i = 0; i < ntemp; ++i) {
            or_mc_propose(seed, (uint32_t)i, OR_TAG_TEMP, L, st, coeff, step, ld, b->LMAX, KMAX);
            es[i] = or_energy_of(b, step, ld, c3ref, c3buf, nthreads);
            s1 += es[i];
        }
        double mean = s1 / ntemp;
        for (int i = 0; i < ntemp; ++i) s2 += (es[i] - mean) * (es[i] - mean);
        T = sqrt(s2 / (ntemp - 1)) * initial_temperature_factor;
        free(es);
    }
    out_scalars[0] = E; out_scalars[1] = T;
    for (int i = 1; i <= nsteps; ++i) {
        T *= temperature_decay;
        or_mc_propose(seed, (uint32_t)i, OR_TAG_MOVE, L, st, coeff, step, ld, b->LMAX, KMAX);
        double En = or_energy_of(b, step, ld, c3ref, c3buf, nthreads);
        double dE = En - E;
        double ap = exp(-dE / T);
        double u = or_uniform(seed, (uint32_t)i, OR_TAG_ACC);
        int acc = 0;
        if (dE < 0.0 || (dE > 0.0 && u < ap)) {
            acc = 1;
            for (int q = 0; q < nl; ++q) st[q] *= stepsizefactor;
            E = En;
            memcpy(coeff, step, sizeof(double) * 2 * ld * KMAX);
        } else {
            for (int q = 0; q < nl; ++q) st[q] /= stepsizefactor;
        }
        E_trace[i - 1] = E; Enew_trace[i - 1] = En; acc_trace[i - 1] = acc;
    }
    out_scalars[2] = E; out_scalars[3] = T;
    free(st); free(c3buf); free(step);
    return 0;
}

/* rotate_all_at_once in resumable pieces (determination.jl:281-338):        */
/* begin != 0 starts a run (E, T from calculate_temperature, step sizes      */
/* initial_stepsize (l - 1), step 0); otherwise st = {step, E, T, naccepted} */
/* and stepsizes[] continue a saved run.  Steps until max_steps are done or, */
/* with min_stepsize > 0, `while stepsizes[L] > minstepsize` (:1, :322)      */
/* fails before a step.  Returns the steps taken.                            */
int or_anneal_ex(const or_basis *b, const double *c3ref, double *coeff, int64_t ld, int KMAX, int begin,
                 double *st /* in/out: step, E, T, naccepted, E0, T0 */, double *stepsizes /* in/out [L/2] */,
                 int max_steps, double min_stepsize, double initial_stepsize, double stepsizefactor,
                 double initial_temperature_factor, double temperature_decay, uint64_t seed, int ntemp,
                 double *E_trace, double *Enew_trace, int *acc_trace, int nthreads)
{
    if (b->lambda != 0.0) return -1;
    int L = b->L, nl = L / 2;
    double *c3buf = (double *)malloc(sizeof(double) * b->R * b->klength);
    double *step = (double *)malloc(sizeof(double) * 2 * ld * KMAX);
    if (begin) {
        for (int l = 2; l <= L; l += 2) stepsizes[l / 2 - 1] = initial_stepsize * (l - 1);   /* :295 */
        double E = or_energy_of(b, coeff, ld, c3ref, c3buf, nthreads);
        double T = 0.0;
        if (initial_temperature_factor > 0.0 && ntemp > 1) {
            double s1 = 0, s2 = 0;
            double *es = (double *)malloc(sizeof(double) * ntemp);
            for (int i = 0; i < ntemp; ++i) {
                or_mc_propose(seed, (uint32_t)i, OR_TAG_TEMP, L, stepsizes, coeff, step, ld, b->LMAX, KMAX);
                es[i] = or_energy_of(b, step, ld, c3ref, c3buf, nthreads);
                s1 += es[i];
            }
            double mean = s1 / ntemp;
            for (int i = 0; i < ntemp; ++i) s2 += (es[i] - mean) * (es[i] - mean);
            T = sqrt(s2 / (ntemp - 1)) * initial_temperature_factor;
            free(es);
        }
        st[0] = 0; st[1] = E; st[2] = T; st[3] = 0; st[4] = E; st[5] = T;
    }
    int64_t i0 = (int64_t)st[0];
    double E = st[1], T = st[2];
    int nacc = (int)st[3], taken = 0;
    for (int j = 0; j < max_steps; ++j) {
        if (min_stepsize > 0.0 && !(stepsizes[nl - 1] > min_stepsize)) break;
        uint32_t i = (uint32_t)(i0 + j + 1);
        T *= temperature_decay;
        or_mc_propose(seed, i, OR_TAG_MOVE, L, stepsizes, coeff, step, ld, b->LMAX, KMAX);
        double En = or_energy_of(b, step, ld, c3ref, c3buf, nthreads);
        double dE = En - E;
        double ap = exp(-dE / T);
        double u = or_uniform(seed, i, OR_TAG_ACC);
        int acc = 0;
        if (dE < 0.0 || (dE > 0.0 && u < ap)) {
            acc = 1;
            for (int q = 0; q < nl; ++q) stepsizes[q] *= stepsizefactor;
            E = En;
            memcpy(coeff, step, sizeof(double) * 2 * ld * KMAX);
        } else {
            for (int q = 0; q < nl; ++q) stepsizes[q] /= stepsizefactor;
        }
        nacc += acc;
        E_trace[j] = E; Enew_trace[j] = En; acc_trace[j] = acc;
        ++taken;
    }
    st[0] = (double)(i0 + taken); st[1] = E; st[2] = T; st[3] = nacc;
    free(c3buf); free(step);
    return taken;
}

/* loadHistograms for the triplet histogram (datagen.jl:432-475), as a plain */
/* restatement: NA x NB x Kf^3 (column-major) -> N x 2N x K3^3; doubling of  */
/* an N x N histogram for k1 >= k2 >= k3 (:445-458), clamp to >= 1e-30       */
/* (:460-462), cut to K3 shells, division by the total of |entries|          */
/* (:465-468).                                                               */
void or_load_c3(const double *c3_full, int NA, int NB, int Kf, int K3, double *out)
{
    int64_t N = NA, R = 2 * N * N, Rf = (int64_t)NA * NB;
    for (int k1 = 0; k1 < K3; ++k1)
        for (int k2 = 0; k2 < K3; ++k2)
            for (int k3 = 0; k3 < K3; ++k3)
                for (int64_t b = 0; b < 2 * N; ++b)
                    for (int64_t a = 0; a < N; ++a) {
                        double v = 0.0;
                        if (NB == 2 * NA)
                            v = c3_full[a + N * b + R * (k3 + (int64_t)Kf * (k2 + (int64_t)Kf * k1))];
                        else if (k1 >= k2 && k2 >= k3)
                            v = c3_full[a + N * (b % N) + Rf * (k3 + (int64_t)Kf * (k2 + (int64_t)Kf * k1))];
                        out[a + N * b + R * (k3 + (int64_t)K3 * (k2 + (int64_t)K3 * k1))] = v < 1e-30 ? 1e-30 : v;
                    }
    int64_t n = R * K3 * K3 * K3;
    double t = 0.0;
    for (int64_t i = 0; i < n; ++i) t += fabs(out[i]);
    for (int64_t i = 0; i < n; ++i) out[i] /= t;
}

/* ------------------------------------------------------------------------ */
/* Wigner-D: d_work / Dlms / rotationMatrix (spherical_harmonics.jl:572-595),*/
/* explicit sum with factorials (long double in place of BigInt).            */
/* out: (2l+1)^2 complex row-major [m+l][s+l].                                */
/* ------------------------------------------------------------------------ */
double or_d_work(int l, int m1, int m2, double beta)
{
    or_fact_init();
    int t1 = l + m1, t2 = l - m2, t3 = m2 - m1;
    int vmin = (-t3 < 0) ? 0 : -t3;
    int vmax = (t2 < t1) ? t2 : t1;
    long double factor = sqrtl(or_fact_ld[l + m1] * or_fact_ld[l - m1] * or_fact_ld[l + m2] * or_fact_ld[l - m2]);
    long double val = 0;
    long double cb = cosl((long double)beta / 2), sb = sinl((long double)beta / 2);
    for (int s = vmin; s <= vmax; ++s) {
        long double t = powl(cb, 2 * l + m1 - m2 - 2 * s) * powl(sb, m2 - m1 + 2 * s);
        if ((s + m2 - m1) & 1) t = -t;
        t /= or_fact_ld[l + m1 - s] * or_fact_ld[s] * or_fact_ld[m2 - m1 + s] * or_fact_ld[l - m2 - s];
        val += t;
    }
    return (double)(val * factor);
}

void or_rotation_matrix(double theta, double phi, double gamma, int l, double *out)
{
    int n = 2 * l + 1;
    for (int m = -l; m <= l; ++m)
        for (int s = -l; s <= l; ++s) {
            double d = or_d_work(l, m, s, theta);
            double ang = -s * gamma - m * phi;
            out[2 * ((m + l) * n + (s + l))] = d * cos(ang);
            out[2 * ((m + l) * n + (s + l)) + 1] = d * sin(ang);
        }
}

/* ------------------------------------------------------------------------ */
/* Synthetic input: point atoms of unit form factor, I(q) = |sum e^{i q.r}|^2 */
/* on shells q_k = k dq (k = 1..KMAX), expanded in Y_lm (Ylm = sphPlm e^{im  */
/* phi}, spherical_harmonics.jl:89-91) by Gauss-Legendre (cos theta) x       */
/* uniform-phi quadrature: c_lm = sum w_a (2pi/nphi) I(a,b) conj(Y_lm(a,b)).  */
/* Coefficients with l <= lfill are written (seanindex layout), the rest 0.  */
/* This is the input generator that feeds both paths; it is not on the hot   */
/* path.                                                                      */
/* ------------------------------------------------------------------------ */
void or_gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < n; ++i) {
        long double z = cosl((long double)OR_PI * (i + 0.75L) / (n + 0.5L)), pp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) { long double p3 = p2; p2 = p1; p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j; }
            pp = n * (z * p1 - p2) / (z * z - 1);
            long double z1 = z;
            z = z1 - p1 / pp;
            if (fabsl(z - z1) < 1e-19L) break;
        }
        x[i] = (double)z;
        w[i] = (double)(2.0L / ((1 - z * z) * pp * pp));
    }
}

void or_intensity_from_atoms(const double *atoms /* natoms x 3 */, int natoms, int KMAX, double dq, int LMAX, int lfill,
                             int ntheta, int nphi, double *coeff /* complex, LMAX^2 x KMAX */, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    (void)nthreads;
    int64_t ld = (int64_t)LMAX * LMAX;
    memset(coeff, 0, sizeof(double) * 2 * ld * KMAX);
    double *gx = (double *)malloc(sizeof(double) * ntheta), *gw = (double *)malloc(sizeof(double) * ntheta);
    or_gauss_legendre(ntheta, gx, gw);
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 1; k <= KMAX; ++k) {
        double q = k * dq;
        double *I = (double *)malloc(sizeof(double) * ntheta * nphi);
        for (int a = 0; a < ntheta; ++a) {
            double ct = gx[a], st = sqrt((1 - ct) * (1 + ct));
            for (int bph = 0; bph < nphi; ++bph) {
                double ph = 2 * OR_PI * bph / nphi;
                double qx = q * st * cos(ph), qy = q * st * sin(ph), qz = q * ct;
                double re = 0, im = 0;
                for (int j = 0; j < natoms; ++j) {
                    double d = qx * atoms[3 * j] + qy * atoms[3 * j + 1] + qz * atoms[3 * j + 2];
                    re += cos(d); im += sin(d);
                }
                I[a * nphi + bph] = re * re + im * im;
            }
        }
        double *ck = coeff + 2 * (int64_t)(k - 1) * ld;
        for (int l = 0; l <= lfill && l < LMAX; ++l)
            for (int m = -l; m <= l; ++m) {
                double re = 0, im = 0;
                for (int a = 0; a < ntheta; ++a) {
                    double y = or_sphPlm(l, m, acos(gx[a]));
                    double sr = 0, si = 0;
                    for (int bph = 0; bph < nphi; ++bph) {
                        double ph = 2 * OR_PI * bph / nphi;
                        sr += I[a * nphi + bph] * cos(m * ph);
                        si -= I[a * nphi + bph] * sin(m * ph);
                    }
                    re += gw[a] * y * sr; im += gw[a] * y * si;
                }
                int64_t s = or_seanindex(l, m, LMAX) - 1;
                ck[2 * s] = re * 2 * OR_PI / nphi;
                ck[2 * s + 1] = im * 2 * OR_PI / nphi;
            }
        free(I);
    }
    free(gx); free(gw);
}

/* ------------------------------------------------------------------------ */
/* Learned proposer: MLP 192 -> 256 -> 256 -> 256 (ReLU) -> 64, MSE on the   */
/* first 36 outputs.  No counterpart in the reference (the reference        */
/* proposes moves only by random_rotation_step, utilities.jl:180-190); this */
/* is the checker of the device's forward/backward kernels.                 */
/* Parameters flat: W1 [256][192], b1 [256], W2 [256][256], b2, W3, b3,      */
/* W4 [64][256], b4 [64].  emulate != 0 rounds, as the device does, the      */
/* inputs, the weights, every activation after ReLU, dY and every masked     */
/* hidden gradient to bfloat16 (round to nearest even of the float value);  */
/* accumulation is FP64 here (FP32 on the device).                          */
/* Pinned by central finite differences of the non-emulated loss.           */
/* ------------------------------------------------------------------------ */
#define OR_MLP_IN 192
#define OR_MLP_HID 256
#define OR_MLP_OUT 64
#define OR_MLP_Q 36

static double or_bf16(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

int or_mlp_nparam(void)
{
    return OR_MLP_HID * OR_MLP_IN + OR_MLP_HID + 2 * (OR_MLP_HID * OR_MLP_HID + OR_MLP_HID) +
           OR_MLP_OUT * OR_MLP_HID + OR_MLP_OUT;
}

/* loss = (1/B) sum_b sum_{j<36} (y_bj - t_bj)^2; grad (may be NULL) gets
 * d loss / d params; Y (may be NULL) gets the [B][64] outputs. */
double or_mlp_loss_grad(const float *params, int B, const float *X /* [B][192] */, const float *T /* [B][36] */,
                        int emulate, double *grad, double *Y)
{
    const int nin[4] = {OR_MLP_IN, OR_MLP_HID, OR_MLP_HID, OR_MLP_HID};
    const int nout[4] = {OR_MLP_HID, OR_MLP_HID, OR_MLP_HID, OR_MLP_OUT};
    int off[4];
    off[0] = 0;
    for (int L = 1; L < 4; ++L) off[L] = off[L - 1] + nout[L - 1] * nin[L - 1] + nout[L - 1];
    double *W[4], *bvec[4];
    for (int L = 0; L < 4; ++L) {
        W[L] = (double *)malloc(sizeof(double) * nout[L] * nin[L]);
        bvec[L] = (double *)malloc(sizeof(double) * nout[L]);
        for (int i = 0; i < nout[L] * nin[L]; ++i) {
            const double w = params[off[L] + i];
            W[L][i] = emulate ? or_bf16(w) : w;
        }
        for (int i = 0; i < nout[L]; ++i) bvec[L][i] = params[off[L] + nout[L] * nin[L] + i];
    }
    /* activations a[0] = x, a[1..3] hidden (post ReLU), a[4] = y */
    double *a[5];
    a[0] = (double *)malloc(sizeof(double) * B * OR_MLP_IN);
    for (int i = 0; i < B * OR_MLP_IN; ++i) a[0][i] = emulate ? or_bf16(X[i]) : X[i];
    for (int L = 0; L < 4; ++L) {
        a[L + 1] = (double *)malloc(sizeof(double) * B * nout[L]);
        for (int b = 0; b < B; ++b)
            for (int o = 0; o < nout[L]; ++o) {
                double s = 0.0;
                for (int i = 0; i < nin[L]; ++i) s += a[L][b * nin[L] + i] * W[L][o * nin[L] + i];
                s += bvec[L][o];
                if (L < 3) {
                    s = s > 0.0 ? s : 0.0;
                    if (emulate) s = or_bf16(s);
                } else if (emulate) {
                    s = (double)(float)s;
                }
                a[L + 1][b * nout[L] + o] = s;
            }
    }
    double loss = 0.0;
    double *d = (double *)calloc((size_t)B * OR_MLP_OUT, sizeof(double));
    for (int b = 0; b < B; ++b)
        for (int j = 0; j < OR_MLP_Q; ++j) {
            const double e = a[4][b * OR_MLP_OUT + j] - T[b * OR_MLP_Q + j];
            loss += e * e / B;
            const double g = 2.0 * e / B;
            d[b * OR_MLP_OUT + j] = emulate ? or_bf16(g) : g;
        }
    if (Y) memcpy(Y, a[4], sizeof(double) * B * OR_MLP_OUT);
    if (grad) {
        for (int L = 3; L >= 0; --L) {
            double *gW = grad + off[L], *gb = grad + off[L] + nout[L] * nin[L];
            for (int o = 0; o < nout[L]; ++o) {
                double sb = 0.0;
                for (int b = 0; b < B; ++b) sb += d[b * nout[L] + o];
                gb[o] = sb;
                for (int i = 0; i < nin[L]; ++i) {
                    double s = 0.0;
                    for (int b = 0; b < B; ++b) s += d[b * nout[L] + o] * a[L][b * nin[L] + i];
                    gW[o * nin[L] + i] = s;
                }
            }
            if (L == 0) break;
            double *dn = (double *)malloc(sizeof(double) * B * nin[L]);
            for (int b = 0; b < B; ++b)
                for (int i = 0; i < nin[L]; ++i) {
                    double s = 0.0;
                    for (int o = 0; o < nout[L]; ++o) s += d[b * nout[L] + o] * W[L][o * nin[L] + i];
                    if (!(a[L][b * nin[L] + i] > 0.0)) s = 0.0;   /* ReLU derivative of the layer below */
                    dn[b * nin[L] + i] = emulate ? or_bf16(s) : s;
                }
            free(d);
            d = dn;
        }
    }
    free(d);
    for (int L = 0; L < 5; ++L) free(a[L]);
    for (int L = 0; L < 4; ++L) { free(W[L]); free(bvec[L]); }
    return loss;
}

// tp_hip.hip -- libtp: host side and C ABI (include/tp.h) of the gfx950
// implementation of the ThreePhotons.jl three-photon correlation c3 and its
// Bayes energy.  Device code is in tp_kernels.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <random>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/tp.h"
#include "tp_host_math.hpp"
#include "tp_kernels.hpp"
#include "tp_triple16.hpp"
#include "tp_triple_sym.hpp"
#include "tp_hgemm.hpp"
#include "tp_batch.hpp"
#include "tp_mlp.hpp"
#include "tp_grad.hpp"
#include "tp_hist.hpp"
#include "tp_c2.hpp"
#include "tp_retrieve.hpp"
#include "tp_fit.hpp"
#include "tp_phase.hpp"
#include "tp_structure.hpp"
#include "tp_ewald.hpp"
#include "tp_pictures.hpp"

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

template <class T>
static int dev_alloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) return tp_fail(TP_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

static int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

static uint16_t host_f16(double x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
static double host_f16_val(uint16_t h) { return (double)__builtin_bit_cast(_Float16, h); }
static uint16_t host_bf16(double x)   // round to nearest even of the float value
{
    const float f = (float)x;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// ---------------------------------------------------------------------------
// objects
// ---------------------------------------------------------------------------
struct PhTables;

// Device scratch of tp_energy_noalloc, for callers that cannot allocate: one
// array per GPU, held from the library's load until the process ends.  Host
// code never uses the name `scratch` as an address: tp_init takes each
// device's address of it once with hipGetSymbolAddress (g_scratch_ptr).
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;   // 32 MiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

struct tp_ctx {
    int device = 0;
    void *scratch = nullptr;   // this device's address of the static scratch array
    hipStream_t stream = nullptr;
    bool timing = false;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double last_ms[4] = {0, 0, 0, 0};
    double *h_pin = nullptr;   // pinned host landing buffer for scalar results
    double *h_res = nullptr;   // pinned, mapped, coherent host buffer the final kernel writes the energy into
    double *d_res = nullptr;   // its device address
    uint64_t res_seq = 0;      // sequence number of the last synchronous evaluation (h_res[1] as uint64)
    tp_batch *eb = nullptr;    // plan (and graph) of the last mixed-precision tp_energy_batch
    std::vector<const void *> eb_sig;
    std::vector<PhTables *> ph_cache;   // SH / Hankel tables of recent (LMAX, K), newest last
};
static void ph_cache_free(tp_ctx *ctx);

// contraction tiles
constexpr int G1_BM = 64, G1_BN = 32, G1_WGM = 2;   // tile of the gradient's gAs = H^T . gG
constexpr int kNumTiles = 12;                       // contraction candidates of the batch autotuner
constexpr int kDgRB = 3, kDgPF = 4;                 // G = H . As (k_dgemm_g): row blocks per workgroup, k-steps
                                                    // of operand prefetch
constexpr int G2_BM = 64, G2_BN = 64, G2_WGM = 4;   // c3 = Bm . G (+ energy)
constexpr int KL_ALIGN = 128;                       // column padding (multiple of G1_BN, G2_BN and H_BN)
constexpr int kMfmaMaxK = 27;                       // k_triple_mfma: at most 16 row blocks of pairs per half

struct tp_basis {
    tp_ctx *ctx = nullptr;
    int L = 0, LMAX = 0, N = 0, K = 0, for_intensity = 1, lstep = 2;
    double lambda = 0, dq = 0;
    int64_t basislen = 0, n_lt = 0, klength = 0, R = 0;
    int n_c = 0;                 // compact coefficients per shell (l in lrange, m = -l..l)
    std::vector<int> coff;       // compact offset of order l (indexed by l)
    std::vector<int> sean0;      // [n_c] seanindex-1 of each compact entry
    int n_m = 0;                 // folded (m2,mu3) classes (even m only when lambda dq = 0)
    // curved Ewald sphere (lambda dq != 0): G formed by k_glam
    bool lam = false;
    int lam_yrow = 0;
    double *d_lamY = nullptr;      // [(L+1)(2L+1)][K] sphPlm(l, m, theta_k)
    int *d_ktk = nullptr;          // [klength] k1 | k2 << 10 | k3 << 20
    LamRec *d_lament = nullptr;    // entries of every class, class-major
    int *d_lamcs = nullptr;        // [n_m + 1] first entry of each class
    int *d_lamyr = nullptr;        // sphPlm rows staged per class: (l, m1), (l, m2), (l, mu3) for every l
    int *d_lamys = nullptr;        // [n_m + 1] first staged row of each class
    int lam_rows = 0;              // most rows staged for one class
    size_t lam_lds = 0;            // k_glam's dynamic LDS
    // the gradient: the same terms grouped by l-triple (dE/dAs = H_kt^T dE/dG)
    LamRec *d_lambent = nullptr;   // entries of every l-triple, l-triple-major
    int *d_lambcs = nullptr;       // [n_lt + 1] first entry of each l-triple
    int *d_lambyr = nullptr;       // sphPlm rows staged per l-triple: (l1, m), (l2, m), (l3, mu) for every m
    int *d_lambys = nullptr;       // [n_lt + 1] first staged row of each l-triple
    int lamb_rows = 0;
    size_t lamb_lds = 0;
    double *d_gG = nullptr;        // [n_m][kl_pad] dE/dG (the sum of the backward contraction's row blocks)
    int lam_nt = 0;                // k_glam's workgroup size (columns per workgroup)
    int64_t R_pad = 0, nlt_pad = 0, kl_pad = 0, nm_pad64 = 0, nm_pad16 = 0;
    int n_rt = 0, n_ct = 0;
    int n_groups = 0, nchunks = 0, Gmax = 0, Emax = 0, npairs = 0, PB = 0;
    size_t triple_lds = 0;
    // MFMA triple products (K <= kMfmaMaxK): work list of (lt, k1 half)
    bool use_mfma = false;
    int nwork = 0, Lw = 0, EmaxLt = 0, Prows = 0;
    size_t mfma_lds = 0, t16_lds = 0;   // dynamic LDS of k_triple_mfma, k_triple16
    int EmaxLt8 = 0;                    // entries of an l-triple with every group padded to a multiple of 8
    float *d_ew8 = nullptr;             // [..] weights, groups padded to multiples of 8 with zeros
    uint32_t *d_eidx8 = nullptr;        // [..]
    int *d_gstart8 = nullptr;           // [n_groups + 1]
    double *d_shsc = nullptr;           // [2K] shell scales of the mixed-precision triple products, shell norms
    int *d_work = nullptr;         // [nwork] lt | half << 16, most expensive first
    // sorted l-triples (k_triple_sym; the library's own weights, K <= 32)
    bool use_sym = false;
    int nsym_work = 0;
    size_t sym_lds = 0, sym16_lds = 0, sym_lds_b = 0;   // dynamic LDS of k_triple_sym, k_triple_sym16, k_triple_sym_batch
    SymItem *d_sym_work = nullptr; // [nsym_work] work items, most expensive first
    int *d_perms = nullptr;        // [n_lt][8] orderings of each sorted l-triple
    int *d_gorder = nullptr;       // [n_groups] run order of the mu3 groups of sorted l-triples
    uint32_t *d_symtab = nullptr;  // V entries each permutation stores (k_triple_sym)
    int *d_symoff = nullptr;       // [6 * 2 * 2 * 2 + 1]
    std::vector<std::pair<int, double>> h_sym_cost;   // per sorted l-triple: (ts, stage-1 + write cost)
    std::vector<SymItem> h_sym_item;                  // per l-triple: its work item with h1 = h2 = 0
    std::vector<int> h_lt_cnt, h_lt_l3;   // per l-triple: basis entries, l3 (work lists of batches)
    // mixed-precision contraction (TP_PREC_FP16X3 / TP_PREC_BF16)
    int prec = TP_PREC_FP64;
    std::vector<double> h_H, h_Bm; // host copies [nlt_pad][nm_pad64], [nm_pad16][R_pad] for building F
    int ldk = 0, n_rt16 = 0, n_ct16 = 0;
    int64_t R_pad16 = 0;
    double F_inv = 1.0;            // 1 / (power-of-two scale of F)
    uint16_t *d_Fhi = nullptr, *d_Flo = nullptr, *d_Fbf = nullptr;   // [R_pad16][ldk]
    uint16_t *d_Ahi = nullptr, *d_Alo = nullptr;                      // [kl_pad][ldk]
    double *d_Ainv = nullptr;      // [kl_pad] 1 / column scale of As
    // factored mixed-precision contraction (TP_PREC_FP16X3_T64): G = H . As on
    // FP16 hi/lo (k_hgemm_g), then the FP64 separated grid
    uint16_t *d_Hhi = nullptr, *d_Hlo = nullptr;   // [kHgRows][ldh16]  H[class][lt] scaled by 1 / H_inv
    int ldh16 = 0;                 // n_lt rounded up to 128 (the K of k_hgemm_g)
    double H_inv = 1.0;
    int nrb16 = 0;                 // row blocks of 16 classes
    double *d_nrm2 = nullptr;      // [K][kNrmLd] l-block norms of the evaluated shells
    // device tables
    double *d_H = nullptr;         // [nlt_pad][nm_pad64]   H[lt][class]
    double *d_Bm = nullptr;        // [nm_pad16][R_pad]     Bm[class][r]
    // separated grid (k_c3sep): classes of one m2 contiguous
    int n_m2 = 0, jpad = 0, sep_kb = 0, njc = 0;
    bool use_sep = false;          // N <= 32, 2(L/2+1) <= 24 and <= 24 classes per m2
    bool sep_wide = false;         // otherwise the wide k_c3sep: N <= 32, 2(L+1) <= 40 (odd m, lambda dq != 0)
    // energies on the folded grid (k_c3sep FOLD): lambda dq = 0 (even m2, mu3)
    // and even N; njf j blocks of 16 cover its N values of j
    bool fold = false;
    int njf = 0;
    int *d_m2start = nullptr;      // [n_m2 + 1]
    double *d_TA = nullptr;        // [2][n_m][jpad]  cos / sin(mu3 b_j)
    double *d_TB = nullptr;        // [round_up(N,16)][sep_kb]  cos(m2 a_i), -sin(m2 a_i)
    double *d_logtab = nullptr;    // [kLogTabSize] tables of tp_fastlog
    double *d_ew = nullptr;        // [basislen] weights, sorted (lt, mu3, m1)
    uint32_t *d_eidx = nullptr;    // [basislen] compact i1 | i2 << 16
    int *d_gstart = nullptr;       // [n_groups + 1]
    int4 *d_ltinfo = nullptr;      // [n_lt] {coff[l3], first group, l3, 0}
    int4 *d_chunks = nullptr;      // [nchunks] {lt0, lt1, g0, g1}
    int *d_pairs = nullptr;        // [npairs] k1 | k2 << 16 (1-based)
    int *d_sean0 = nullptr;        // [n_c]
    int *d_coff = nullptr;         // [L+1]
    // per-evaluation scratch
    double *d_As = nullptr;        // [nlt_pad][kl_pad]
    double *d_G = nullptr;         // [nm_pad64][kl_pad]
    double *d_part = nullptr;      // [n_rt][2][kl_pad]
    double *d_out = nullptr;       // [8] scalar results
    double2 *d_raw = nullptr;      // [K][LMAX^2] staging of host coefficients
    double2 *d_coef = nullptr;     // [K][n_c] compact coefficients of host-input calls
    double *d_c3 = nullptr;        // [klength][R] c3raw output staging (lazy)
    // energy gradient (tp_energy_grad), built on first use
    std::vector<int4> h_ltinfo;    // host copy of d_ltinfo
    bool grad_ready = false;
    int nlt_pad64 = 0, gLw = 0, gEmax = 0;
    size_t bwd_lds = 0;
    double *d_HT4 = nullptr;       // [njc * nm_pad64][nlt_pad64]  H^T stacked njc times
    double *d_gGp = nullptr;       // [njc * nm_pad64][kl_pad]     dE/dG per block of 16 values of j
    double *d_gA = nullptr;        // [nlt_pad64][kl_pad]          dE/dAs
    double2 *d_gpart = nullptr;    // [n_lt][3][K][gLw]            per (l-triple, slot) gradient blocks
    double2 *d_grad = nullptr;     // [K][n_c]                     dE/dRe c + i dE/dIm c
    int *d_lstart = nullptr;       // [L/2 + 2]
    int *d_llist = nullptr;        // [3 n_lt] lt * 3 + slot, grouped by the slot's l
    int *d_bwd_order = nullptr;    // [n_lt] l-triples by decreasing backward cost
    // two-photon correlation (tp_two_photons), grown on demand
    int64_t c2_shells = 0, c2_len = 0;
    double2 *d_c2coef = nullptr;   // [c2_shells][n_c]
    double2 *d_c2raw = nullptr;    // [c2_shells][LMAX^2]
    double *d_c2 = nullptr;        // [c2_len] = N K2 K2
};

struct tp_ref {
    const tp_basis *b = nullptr;
    int klo = 1, khi = 1;      // K3_range of the walk
    int64_t ncols = 0;
    double *d_qhat = nullptr;  // [ncols][R]
    double *d_qfold = nullptr; // [ncols][N^2/2] the histogram on the folded grid (bases with fold)
    double *d_qsum = nullptr;  // [ncols]
    double *d_tf = nullptr;    // [ncols]
};

struct tp_state {
    const tp_basis *b = nullptr;
    int nshell = 0;
    double2 *d_cur = nullptr;    // [nshell][n_c] the current structure between calls
    double2 *d_prop = nullptr;   // [nshell][n_c] proposal buffer
    double *d_scal = nullptr;    // [8]: E, T
    double *d_step = nullptr;    // [L/2] step sizes
    int *d_sel = nullptr;        // [4] which buffer is current during tp_anneal, then the McArgs
                                 // control block {stopped, last step, accepted steps}
};

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
static int check_ctx(tp_ctx *ctx)
{
    if (!ctx) return tp_fail(TP_E_ARG, "null context");
    TP_HIP(hipSetDevice(ctx->device));
    return 0;
}

// Upload host coefficients of `nshell` shells and pack them to the compact
// layout in dst.
static int upload_coeff(tp_ctx *ctx, tp_basis *b, const double *coeff, int64_t ld, int nshell, double2 *dst,
                        double2 *raw_buf)
{
    if (!coeff || ld < (int64_t)b->LMAX * b->LMAX) return tp_fail(TP_E_ARG, "coefficient vector shorter than LMAX^2");
    double2 *raw = raw_buf;
    const bool tmp = !raw || nshell > b->K || ld != (int64_t)b->LMAX * b->LMAX;
    if (tmp)
        if (int rc = dev_alloc(&raw, (size_t)nshell * ld)) return rc;
    TP_HIP(hipMemcpyAsync(raw, coeff, sizeof(double2) * (size_t)nshell * ld, hipMemcpyHostToDevice, ctx->stream));
    const int64_t n = (int64_t)nshell * b->n_c;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, raw, ld, nshell,
                       b->d_sean0, b->n_c, dst);
    TP_HIP(hipGetLastError());
    if (tmp) {
        TP_HIP(hipStreamSynchronize(ctx->stream));
        TP_HIP(hipFree(raw));
    }
    return 0;
}

// Per-evaluation device scratch of enqueue_eval (by default the basis's own
// buffers).
struct EvalScratch {
    double *As, *G, *part, *Ainv, *shsc;
    uint16_t *Ahi, *Alo;
    double *nrm2;   // [K][kNrmLd] l-block norms (k_triple_sym -> k_hgemm_g)
};

static EvalScratch basis_scratch(const tp_basis *b)
{
    return EvalScratch{b->d_As, b->d_G, b->d_part, b->d_Ainv, b->d_shsc, b->d_Ahi, b->d_Alo, b->d_nrm2};
}

// Enqueue one evaluation on coefficients (b0, b1, sel, flip): As, G, c3 (+energy).
// With r == nullptr the raw c3 goes to c3out; otherwise the energy is formed
// and k_final applies `mc`.
static int enqueue_eval(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, const double2 *b0, const double2 *b1,
                        const int *sel, int flip, double *c3out, double *Eout, const McArgs &mc, bool force64 = false,
                        const EvalScratch *scr = nullptr)
{
    const EvalScratch sc = scr ? *scr : basis_scratch(b);
    const bool timing = ctx->timing;
    hipStream_t st = ctx->stream;
    const int prec = force64 ? TP_PREC_FP64 : b->prec;
    if (timing) TP_HIP(hipEventRecord(ctx->ev[0], st));
    const bool c16 = prec == TP_PREC_FP16X3 || prec == TP_PREC_BF16;   // dense FP16 contraction F . As
    // compensated FP16 triple products (k_triple16): TRIPLE16, and the dense
    // modes on bases without the sorted l-triple tables (which otherwise take
    // the mixed sorted-triple kernel, as FP16X3_T64 and their batches do)
    const bool t16 = b->use_mfma && (prec == TP_PREC_TRIPLE16 || (c16 && !b->use_sym));
    if (t16 || c16)   // shell scales (of the FP16 triple products) and norms (the contraction's column scales)
        hipLaunchKernelGGL(k_shell_scales, dim3((unsigned)b->K), dim3(256), 0, st, b0, b1, sel, flip, b->n_c, b->L,
                           sc.shsc);
    if (t16) {
        // mixed precision: compensated FP16 triple products
        hipLaunchKernelGGL(k_triple16, dim3((unsigned)b->nwork), dim3(512), b->t16_lds, st, b0, b1, sel, flip,
                           sc.shsc, b->n_c, b->K, b->d_ew8, b->d_eidx8, b->d_gstart8, b->d_ltinfo, b->d_work, b->Lw,
                           b->EmaxLt8, b->Prows, sc.As, b->kl_pad);
    } else if (b->use_sym) {   // FP64; the mixed modes with stage 1 on FP32 and stage 2 on FP16 MFMA
        if (prec != TP_PREC_FP64)
            hipLaunchKernelGGL(k_triple_sym16, dim3((unsigned)b->nsym_work), dim3(kSymThreads), b->sym16_lds, st, b0, b1,
                               sel, flip, b->n_c, b->K, b->d_ew, b->d_eidx, b->d_gstart, b->d_sym_work,
                               b->d_perms, b->d_gorder, b->d_symtab, b->d_symoff, b->Lw, sc.As, b->kl_pad,
                               b->lam ? nullptr : sc.nrm2);   // norms: column scales of k_hgemm_g
        else
            hipLaunchKernelGGL(k_triple_sym, dim3((unsigned)b->nsym_work), dim3(kSymThreads), b->sym_lds, st, b0, b1,
                               sel, flip, b->n_c, b->K, b->d_ew, b->d_eidx, b->d_gstart, b->d_sym_work,
                               b->d_perms, b->d_gorder, b->d_symtab, b->d_symoff, b->Lw, sc.As, b->kl_pad, nullptr);
    } else if (b->use_mfma) {
        hipLaunchKernelGGL(k_triple_mfma, dim3((unsigned)b->nwork), dim3(1024), b->mfma_lds, st, b0, b1, sel, flip,
                           b->n_c, b->K, b->d_ew, b->d_eidx, b->d_gstart, b->d_ltinfo, b->d_work, b->Lw, b->EmaxLt,
                           b->Prows, sc.As, b->kl_pad);
    } else {
        hipLaunchKernelGGL(k_triple, dim3((unsigned)b->nchunks), dim3(1024), b->triple_lds, st, b0, b1, sel, flip,
                           b->n_c, b->K, b->d_ew, b->d_eidx, b->d_gstart, b->d_ltinfo, b->d_chunks, b->d_pairs,
                           b->npairs, b->PB, b->Gmax, b->Emax, sc.As, b->kl_pad);
    }
    TP_HIP(hipGetLastError());
    if (timing) TP_HIP(hipEventRecord(ctx->ev[1], st));
    const int ncols = r ? (int)r->ncols : (int)b->klength;
    int n_rt_fin = b->n_rt;
    if (c16) {
        // mixed precision: c3 = F . As with 16-bit MFMA operands
        const double *qh = r ? r->d_qhat : nullptr;
        const dim3 grid16((unsigned)(b->n_rt16 * b->n_ct16));
        const dim3 gs((unsigned)(b->kl_pad / 64), (unsigned)(b->ldk / 64));
        if (prec != TP_PREC_BF16) {
            hipLaunchKernelGGL(k_split_as<1>, gs, dim3(256), 0, st, sc.As, b->kl_pad, (int)b->n_lt, b->ldk,
                               sc.shsc + b->K, b->K, (int)b->klength, sc.Ahi, sc.Alo, sc.Ainv);
        } else {
            hipLaunchKernelGGL(k_split_as<2>, gs, dim3(256), 0, st, sc.As, b->kl_pad, (int)b->n_lt, b->ldk,
                               sc.shsc + b->K, b->K, (int)b->klength, sc.Ahi, nullptr, sc.Ainv);
        }
        TP_HIP(hipGetLastError());
        if (timing) TP_HIP(hipEventRecord(ctx->ev[2], st));
        const uint16_t *Fh = prec != TP_PREC_BF16 ? b->d_Fhi : b->d_Fbf;
#define TP_G16(MODE, OUT, EN, OUTP)                                                                                \
    hipLaunchKernelGGL((k_gemm16<H_BM, H_BN, MODE, OUT, EN, 0>), grid16, dim3(256), 0, st, Fh, b->d_Flo,          \
                       (int64_t)b->ldk, sc.Ahi, sc.Alo, (int64_t)b->ldk, b->ldk, b->F_inv, sc.Ainv, (int)b->R,  \
                       ncols, b->n_rt16, b->n_ct16, OUTP, b->R, qh, sc.part, b->kl_pad, nullptr, nullptr)
        if (prec != TP_PREC_BF16) {
            if (r && c3out) TP_G16(1, 1, true, c3out);
            else if (r) TP_G16(1, 0, true, nullptr);
            else TP_G16(1, 1, false, c3out);
        } else {
            if (r && c3out) TP_G16(2, 1, true, c3out);
            else if (r) TP_G16(2, 0, true, nullptr);
            else TP_G16(2, 1, false, c3out);
        }
#undef TP_G16
        TP_HIP(hipGetLastError());
        n_rt_fin = b->n_rt16;
    } else {
    if (b->lam) {
        // curved Ewald sphere: the class weights depend on the k-triple
        hipLaunchKernelGGL(k_glam, dim3((unsigned)((b->kl_pad + b->lam_nt - 1) / b->lam_nt), (unsigned)b->n_m),
                           dim3((unsigned)b->lam_nt), b->lam_lds, st, sc.As, b->kl_pad, b->d_ktk, (int)b->klength,
                           b->d_lamY, b->K, b->d_lamyr, b->d_lamys, b->lam_rows, b->d_lament, b->d_lamcs, sc.G,
                           b->kl_pad,
                           nullptr, nullptr);
    } else if (prec == TP_PREC_FP16X3_T64) {
    // G = H . As on FP16 hi/lo MFMA operands (As split while staged), FP64 out
    // (16 columns x 4 row blocks per workgroup; rows past the classes up to
    // nm_pad16 written as zeros for the dense Bm . G form)
    hipLaunchKernelGGL((k_hgemm_g<16, 4, true, 128>), dim3((unsigned)(b->kl_pad / 16), (unsigned)((b->nrb16 + 3) / 4)),
                       dim3(256), 0, st, b->d_Hhi, b->d_Hlo, b->ldh16, b->H_inv, sc.As, b->kl_pad, (int)b->n_lt,
                       b->ldh16, nullptr, sc.nrm2, b->L, b->K, (int)b->nm_pad16, sc.G, b->kl_pad);
    } else {
    // G[class][kt] = sum_lt H[lt][class] As[lt][kt]: 16 columns x 48 classes per
    // workgroup, K split over its waves (rows up to nm_pad16 written; those
    // past the classes are zero)
    const int ncb = (int)(b->kl_pad / 16), nrt = (int)((b->nm_pad16 + 16 * kDgRB - 1) / (16 * kDgRB));
    hipLaunchKernelGGL((k_dgemm_g<kDgRB, kDgPF>), dim3((unsigned)(8 * ((ncb + 7) / 8) * nrt)), dim3(256), 0, st,
                       b->d_H, b->nm_pad64, (int)b->nm_pad64, sc.As, b->kl_pad, (int)b->nlt_pad, ncb, nrt, sc.G,
                       b->kl_pad, (int)b->nm_pad64);
    }
    TP_HIP(hipGetLastError());
    if (timing) TP_HIP(hipEventRecord(ctx->ev[2], st));
    if (!b->use_sep && !b->sep_wide) {
        // c3 = Bm . G as a GEMM (k_c3sep keeps at most 32 values of alpha per
        // workgroup and at most 24 classes per m2 and 24 rows of (m2, cos/sin))
        const dim3 grid((unsigned)(b->n_rt * b->n_ct));
        if (r && c3out)
            hipLaunchKernelGGL((k_gemm<G2_BM, G2_BN, G2_WGM, 1, true>), grid, dim3(256), 0, st, b->d_Bm, b->R_pad,
                               sc.G, b->kl_pad, (int)b->nm_pad16, (int)b->R, ncols, b->n_rt, b->n_ct, c3out, b->R,
                               r->d_qhat, sc.part, b->kl_pad, b->d_logtab);
        else if (r)
            hipLaunchKernelGGL((k_gemm<G2_BM, G2_BN, G2_WGM, 0, true>), grid, dim3(256), 0, st, b->d_Bm, b->R_pad,
                               sc.G, b->kl_pad, (int)b->nm_pad16, (int)b->R, ncols, b->n_rt, b->n_ct, nullptr, b->R,
                               r->d_qhat, sc.part, b->kl_pad, b->d_logtab);
        else
            hipLaunchKernelGGL((k_gemm<G2_BM, G2_BN, G2_WGM, 1, false>), grid, dim3(256), 0, st, b->d_Bm, b->R_pad,
                               sc.G, b->kl_pad, (int)b->nm_pad16, (int)b->R, ncols, b->n_rt, b->n_ct, c3out, b->R,
                               nullptr, nullptr, (int64_t)0, nullptr);
        TP_HIP(hipGetLastError());
        n_rt_fin = b->n_rt;
    } else {
    // c3 = Bm . G on the separated grid, energy fused
    const dim3 grid((unsigned)((ncols + 15) / 16), (unsigned)b->njc);
    const int64_t tap = (int64_t)b->n_m * b->jpad;
#define TP_SEP(OUT, EN, OUTP, QH)                                                                                   \
    hipLaunchKernelGGL((k_c3sep<OUT, EN>), grid, dim3(512), c3sep_lds_bytes(b->sep_kb), st, sc.G, b->kl_pad, b->d_m2start, b->n_m2, b->d_TA, \
                       tap, b->jpad, b->d_TB, b->sep_kb, b->N, ncols, QH, sc.part, b->kl_pad, OUTP, b->R,       \
                       b->d_logtab)
    if (b->sep_wide) {
        // odd-m classes (curved Ewald sphere): the wide instance
        if (r && c3out) hipLaunchKernelGGL((k_c3sep<1, true, false, false, kSepWideQB>), grid, dim3(512),
                                           c3sep_lds_bytes(b->sep_kb), st, sc.G, b->kl_pad, b->d_m2start, b->n_m2,
                                           b->d_TA, tap, b->jpad, b->d_TB, b->sep_kb, b->N, ncols, r->d_qhat, sc.part,
                                           b->kl_pad, c3out, b->R, b->d_logtab);
        else if (r) hipLaunchKernelGGL((k_c3sep<0, true, false, false, kSepWideQB>), grid, dim3(512),
                                       c3sep_lds_bytes(b->sep_kb), st, sc.G, b->kl_pad, b->d_m2start, b->n_m2, b->d_TA,
                                       tap, b->jpad, b->d_TB, b->sep_kb, b->N, ncols, r->d_qhat, sc.part, b->kl_pad,
                                       nullptr, b->R, b->d_logtab);
        else hipLaunchKernelGGL((k_c3sep<1, false, false, false, kSepWideQB>), grid, dim3(512),
                                c3sep_lds_bytes(b->sep_kb), st, sc.G, b->kl_pad, b->d_m2start, b->n_m2, b->d_TA, tap,
                                b->jpad, b->d_TB, b->sep_kb, b->N, ncols, nullptr, sc.part, b->kl_pad, c3out, b->R,
                                b->d_logtab);
        n_rt_fin = b->njc;
    } else if (r && !c3out && b->fold && r->d_qfold) {
        // the energy on the folded grid: njf j blocks
        hipLaunchKernelGGL((k_c3sep<0, true, false, true>), dim3(grid.x, (unsigned)b->njf), dim3(512),
                           c3sep_lds_bytes(b->sep_kb), st, sc.G, b->kl_pad, b->d_m2start, b->n_m2, b->d_TA, tap,
                           b->jpad, b->d_TB, b->sep_kb, b->N, ncols, r->d_qfold, sc.part, b->kl_pad, nullptr, b->R,
                           b->d_logtab);
        n_rt_fin = b->njf;
    } else {
        if (r && c3out) TP_SEP(1, true, c3out, r->d_qhat);
        else if (r) TP_SEP(0, true, nullptr, r->d_qhat);
        else TP_SEP(1, false, c3out, nullptr);
        n_rt_fin = b->njc;
    }
#undef TP_SEP
    TP_HIP(hipGetLastError());
    }
    }
    if (timing) TP_HIP(hipEventRecord(ctx->ev[3], st));
    if (r) {
        hipLaunchKernelGGL(k_final, dim3(1), dim3(kFinalThreads), 0, st, sc.part, n_rt_fin, b->kl_pad, ncols,
                           r->d_qsum, r->d_tf, Eout, mc);
        TP_HIP(hipGetLastError());
    }
    if (timing) TP_HIP(hipEventRecord(ctx->ev[4], st));
    return 0;
}

static McArgs mc_none()
{
    McArgs m;
    std::memset(&m, 0, sizeof(m));
    return m;
}

// Wait for the synchronous evaluation numbered seq: k_final stores the energy
// in mapped host memory and then, with a system-scope release, the sequence
// number after it, which the host polls (a few microseconds sooner than the
// stream's completion signal).  Past a time limit, or with timing events
// recorded, the stream is synchronised instead (which also reports errors).
static int wait_result(tp_ctx *ctx, uint64_t seq)
{
    if (!ctx->timing) {
        const volatile uint64_t *hp = reinterpret_cast<const volatile uint64_t *>(ctx->h_res + 1);
        const auto t0 = std::chrono::steady_clock::now();
        for (uint64_t spin = 0;; ++spin) {
            if (__atomic_load_n(const_cast<const uint64_t *>(hp), __ATOMIC_ACQUIRE) == seq) return 0;
            if ((spin & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(200)) break;
        }
    }
    TP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

static void record_timing(tp_ctx *ctx)
{
    if (!ctx->timing) return;
    for (int i = 0; i < 4; ++i) {
        float t = 0;
        if (hipEventElapsedTime(&t, ctx->ev[i], ctx->ev[i + 1]) == hipSuccess) ctx->last_ms[i] = t;
    }
}

// ---------------------------------------------------------------------------
// batches: many evaluations (uneven K, own histograms) scored by one fixed
// sequence of launches, captured as a HIP graph
// ---------------------------------------------------------------------------
struct tp_batch {
    tp_ctx *ctx = nullptr;
    const tp_basis *b = nullptr;   // tables and F (same L, LMAX, N as every evaluation; K >= theirs)
    int ngroups = 0, nprop = 0, ne = 0, nl = 0, hist_cap = 0;
    bool sweep = false;            // annealing runs (rotation proposals) vs plain evaluations
    std::vector<EvalDesc> h_ev;
    std::vector<GroupState> h_gs;
    std::vector<std::vector<std::pair<double, int2>>> items;   // per evaluation: (cost, {e, lt | h << 16})
    int64_t ncb = 0, ldA = 0, cstride = 0;
    int nwork = 0;
    int prec = TP_PREC_FP16X3;             // the basis's precision mode when the batch was made
    int gcand = 0, order = 0;              // contraction candidate (kTiles), triple-product work order (autotuned)
    bool fold = false;                     // TP_PREC_FP16X3_T64 energies on the folded grid (k_c3sep FOLD)
    bool sym = false;                      // triple products on the sorted l-triples (k_triple_sym_batch, As in FP32)
    int sym_order = 0;                     // work order of the sorted l-triple items (batch_sym_work)
    int family = TP_MOVE_PLANE;            // move family of the random slots
    double tune_ms[2 + kNumTiles] = {};
    EvalDesc *d_ev = nullptr;
    GroupState *d_gs = nullptr;
    int2 *d_work = nullptr;
    SymItem *d_swork = nullptr;    // sorted l-triple work items of every evaluation, by cost
    int nswork = 0;
    double2 *d_cur = nullptr, *d_prop = nullptr;
    double *d_quat = nullptr, *d_As = nullptr, *d_sc = nullptr, *d_isc = nullptr, *d_part = nullptr;
    float *d_As32 = nullptr;       // FP16X3_T64: As in FP32 [n_lt][ldA] (k_triple_sym_batch -> k_hgemm_g)
    double *d_E = nullptr, *d_Ehist = nullptr, *d_shsc = nullptr;   // d_shsc [ne][K] shell scales
    uint16_t *d_Ahi = nullptr, *d_Alo = nullptr;   // dense F . As (TP_PREC_FP16X3, TP_PREC_BF16)
    double *d_G = nullptr;                         // [16 nrb16][ldA] G = H . As (TP_PREC_FP16X3_T64)
    const double **d_qtab = nullptr;
    int *d_nvalid = nullptr;
    int *d_bkt0 = nullptr;         // curved Ewald sphere: first k-triple of each 64-column block
    double2 *d_raw = nullptr;      // staging of host coefficients (plain evaluations)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double g_decay = 0.0, g_factor = 0.0;   // what the captured sweep graph was built with
    std::vector<const tp_ref *> refs;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // move log (training data of the learned proposer)
    int log_cap = 0;
    uint16_t *d_Xlog = nullptr;    // [log_cap][192]
    float *d_Tlog = nullptr;       // [log_cap][36]
    int *d_logn = nullptr;         // [1] samples logged
    // learned proposals: the last nlearn slots of every run
    struct tp_mlp *mlp = nullptr;
    int nlearn = 0, Mi = 0;
    double noise = 0.0;
    uint16_t *d_Xi = nullptr, *d_Hi[3] = {nullptr, nullptr, nullptr};
    float *d_Yi = nullptr;
    // given rotations (tp_batch_eval_rotations), allocated on first use
    double *d_gen = nullptr;       // [ngroups][sum_l (2l+1)^2] generators
    double *d_theta = nullptr, *d_uvb = nullptr;
};

static void batch_free(tp_batch *B)
{
    if (!B) return;
    if (B->ctx) (void)hipSetDevice(B->ctx->device);
    void *geo[] = {B->d_gen, B->d_theta, B->d_uvb};
    for (void *p : geo)
        if (p) (void)hipFree(p);
    B->d_gen = B->d_theta = B->d_uvb = nullptr;
    if (B->exec) (void)hipGraphExecDestroy(B->exec);
    if (B->graph) (void)hipGraphDestroy(B->graph);
    for (auto &e : B->ev)
        if (e) (void)hipEventDestroy(e);
    void *ptrs[] = {B->d_ev,  B->d_gs,  B->d_work, B->d_swork, B->d_cur,  B->d_prop,   B->d_quat,   B->d_As, B->d_As32,
                    B->d_sc,  B->d_isc, B->d_part, B->d_E,    B->d_Ehist,  B->d_Ahi,    B->d_Alo,
                    (void *)B->d_qtab, B->d_nvalid, B->d_bkt0, B->d_raw, B->d_Xlog, B->d_Tlog, B->d_logn, B->d_shsc,
                    B->d_Xi, B->d_Hi[0], B->d_Hi[1], B->d_Hi[2], B->d_Yi, B->d_G};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete B;
}

// Work items of the MFMA triple products for an evaluation with K shells
// (the per-(l-triple, k1 half) costs of tp_basis_create).
static void batch_items(const tp_basis *b, int e, int K, std::vector<std::pair<double, int2>> &out)
{
    const int nh = K > 16 ? 2 : 1;
    for (int h = 0; h < nh; ++h) {
        const int k1lo = 16 * h + 1, k1hi = std::min(16 * h + 16, K);
        const int P = k1hi * (k1hi + 1) / 2 - (k1lo - 1) * k1lo / 2;
        const int nrb = (P + 15) / 16;
        for (int t = 0; t < (int)b->n_lt; ++t) {
            const int n3 = 2 * b->h_lt_l3[t] + 1;
            const double cost = 2.0 * b->h_lt_cnt[t] + 8.0 * ((n3 + 7) / 8) * nrb;
            out.push_back({cost, make_int2(e, t | (h << 16))});
        }
    }
}

// Work items of the sorted l-triple form for every evaluation (both K1 and K2
// halves where the evaluation has more than 16 shells).  order 0: all items by
// decreasing cost.  order 1: evaluation-major per XCD -- the evaluations are
// dealt to the 8 XCDs (workgroup i runs on XCD i mod 8) balancing their
// cost, each XCD runs its evaluations one after another, an evaluation's
// items heaviest first, so the coefficients of the one or two evaluations in
// flight on an XCD stay in its L2; queues are padded to equal length with
// empty items (e = -1).
static int batch_sym_work(tp_batch *B, int order)
{
    const tp_basis *b = B->b;
    if (!b->use_sym) return 0;
    std::vector<std::vector<std::pair<double, SymItem>>> per((size_t)B->ne);
    std::vector<double> ecost((size_t)B->ne, 0.0);
    for (int e = 0; e < B->ne; ++e) {
        const int K = B->h_ev[e].K, nh = K > 16 ? 2 : 1;
        for (const auto &tc : b->h_sym_cost) {
            const int n3 = 2 * b->h_lt_l3[tc.first] + 1;
            const double cost = tc.second + ((n3 + 1) / 2) * 2.0 * nh;
            const int4 li = b->h_ltinfo[tc.first];
            const bool sym12 = (li.w & 0xff) == ((li.w >> 8) & 0xff);   // items h1 < h2 stored by their mirrors
            for (int h1 = 0; h1 < nh; ++h1)
                for (int h2 = 0; h2 < nh; ++h2)
                    if (!(sym12 && h1 < h2)) {
                        SymItem it = b->h_sym_item[tc.first];
                        it.e = e;
                        it.K = K;
                        it.tsx |= h1 << 11 | h2 << 12;
                        per[(size_t)e].push_back({cost, it});
                        ecost[(size_t)e] += cost;
                    }
        }
        std::stable_sort(per[(size_t)e].begin(), per[(size_t)e].end(),
                         [](const std::pair<double, SymItem> &a, const std::pair<double, SymItem> &c) {
                             return a.first > c.first;
                         });
    }
    std::vector<SymItem> w;
    if (order == 0) {
        std::vector<std::pair<double, SymItem>> all;
        for (auto &v : per) all.insert(all.end(), v.begin(), v.end());
        std::stable_sort(all.begin(), all.end(),
                         [](const std::pair<double, SymItem> &a, const std::pair<double, SymItem> &c) {
                             return a.first > c.first;
                         });
        for (auto &x : all) w.push_back(x.second);
    } else {
        std::vector<int> eo((size_t)B->ne);
        for (int e = 0; e < B->ne; ++e) eo[(size_t)e] = e;
        std::stable_sort(eo.begin(), eo.end(), [&](int a, int c) { return ecost[(size_t)a] > ecost[(size_t)c]; });
        std::vector<std::vector<SymItem>> q(8);
        double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int e : eo) {
            int x = 0;
            for (int y = 1; y < 8; ++y)
                if (load[y] < load[x]) x = y;
            for (auto &it : per[(size_t)e]) q[(size_t)x].push_back(it.second);
            load[x] += ecost[(size_t)e];
        }
        size_t len = 0;
        for (auto &v : q) len = std::max(len, v.size());
        SymItem empty{};
        empty.e = -1;
        for (size_t i = 0; i < len; ++i)
            for (int x = 0; x < 8; ++x) w.push_back(i < q[(size_t)x].size() ? q[(size_t)x][i] : empty);
    }
    if (B->d_swork && (int)w.size() != B->nswork) {
        (void)hipFree(B->d_swork);
        B->d_swork = nullptr;
    }
    if (!B->d_swork)
        if (int rc = dev_alloc(&B->d_swork, w.size())) return rc;
    TP_HIP(hipMemcpy(B->d_swork, w.data(), sizeof(SymItem) * w.size(), hipMemcpyHostToDevice));
    B->nswork = (int)w.size();
    B->sym_order = order;
    return 0;
}

// Upload the work list in the given order: 0 = all items by decreasing cost,
// 1 = evaluation-major (an evaluation's items adjacent, by decreasing cost).
static int batch_upload_work(tp_batch *B, int order)
{
    std::vector<std::pair<double, int2>> all;
    for (auto &v : B->items) all.insert(all.end(), v.begin(), v.end());
    auto cmp_cost = [](const std::pair<double, int2> &a, const std::pair<double, int2> &b) {
        if (a.first != b.first) return a.first > b.first;
        if (a.second.x != b.second.x) return a.second.x < b.second.x;
        return a.second.y < b.second.y;
    };
    auto cmp_eval = [&](const std::pair<double, int2> &a, const std::pair<double, int2> &b) {
        if (a.second.x != b.second.x) return a.second.x < b.second.x;
        return cmp_cost(a, b);
    };
    if (order == 0) std::sort(all.begin(), all.end(), cmp_cost);
    else std::sort(all.begin(), all.end(), cmp_eval);
    std::vector<int2> w(all.size());
    for (size_t i = 0; i < all.size(); ++i) w[i] = all[i].second;
    if (!B->d_work)
        if (int rc = dev_alloc(&B->d_work, w.size())) return rc;
    TP_HIP(hipMemcpy(B->d_work, w.data(), sizeof(int2) * w.size(), hipMemcpyHostToDevice));
    B->nwork = (int)w.size();
    B->order = order;
    return 0;
}

// Lay out ngroups groups of nprop evaluations each; group g has refs[g]
// (its histogram; K = refs[g]->b->K).
static int batch_build(tp_ctx *ctx, const tp_basis *b, int ngroups, int nprop, const tp_ref *const *refs,
                       const int *Kg, bool sweep, int hist_cap, tp_batch **out)
{
    *out = nullptr;
    if (!b->use_mfma) return tp_fail(TP_E_UNSUPPORTED, "batches need the MFMA triple products (K <= 27)");
    if (b->prec != TP_PREC_FP16X3 && b->prec != TP_PREC_BF16 && b->prec != TP_PREC_FP16X3_T64)
        return tp_fail(TP_E_UNSUPPORTED, "batched contraction needs TP_PREC_FP16X3, TP_PREC_FP16X3_T64 or TP_PREC_BF16");
    const bool fac = b->prec == TP_PREC_FP16X3_T64;   // factored contraction G = H . As, c3 = Bm . G
    if (!fac && b->R % 128 != 0)
        return tp_fail(TP_E_UNSUPPORTED, "batched contraction needs 2N^2 to be a multiple of 128");
    if (fac && (!b->use_sym || !(b->use_sep || (b->lam && b->sep_wide)) || (!b->lam && !b->d_Hhi)))
        return tp_fail(TP_E_UNSUPPORTED, "TP_PREC_FP16X3_T64 batches need the sorted l-triple tables and the "
                                         "separated grid (N <= 32)");
    tp_batch *B = new tp_batch();
    B->ctx = ctx;
    B->b = b;
    B->prec = b->prec;
    B->ngroups = ngroups;
    B->nprop = nprop;
    B->ne = ngroups * nprop;
    B->nl = b->L / 2;
    B->sweep = sweep;
    B->hist_cap = hist_cap;
    B->fold = fac && b->fold;
    B->sym = b->use_sym;
    for (int g = 0; g < ngroups; ++g) B->fold = B->fold && refs[g]->d_qfold;
    B->cstride = (int64_t)b->K * b->n_c;
    B->refs.assign(refs, refs + ngroups);
    // column blocks
    std::vector<const double *> qtab;
    std::vector<int> nvalid, bkt0;
    int64_t cb = 0;
    for (int g = 0; g < ngroups; ++g) {
        const tp_ref *r = refs[g];
        const int K = Kg[g];
        const int64_t kl = (int64_t)K * (K + 1) * (K + 2) / 6, nt = (kl + 63) / 64;
        for (int p = 0; p < nprop; ++p) {
            EvalDesc d;
            std::memset(&d, 0, sizeof(d));
            d.K = K;
            d.ncols = (int)r->ncols;
            d.cbase = (int)cb;
            d.nprop = nprop;
            d.p = p;
            d.group = g;
            d.qsum = r->d_qsum;
            d.tf = r->d_tf;
            B->h_ev.push_back(d);
        }
        for (int64_t j = 0; j < nt; ++j)
            for (int p = 0; p < nprop; ++p) {
                const int64_t nv = std::max<int64_t>(0, std::min<int64_t>(64, r->ncols - j * 64));
                if (B->fold)   // folded columns of N^2/2 values
                    qtab.push_back(r->d_qfold + (nv > 0 ? j * 64 * (b->R / 4) : 0));
                else
                    qtab.push_back(r->d_qhat + (nv > 0 ? j * 64 * b->R : 0));
                nvalid.push_back((int)nv);
                bkt0.push_back((int)(j * 64));
            }
        GroupState s;
        std::memset(&s, 0, sizeof(s));
        s.amax = 1.0;
        s.K = K;
        s.nprop = nprop;
        s.e0 = g * nprop;
        B->h_gs.push_back(s);
        cb += nt * nprop;
    }
    while (cb % 2) {   // even block count: 128-column tiles
        qtab.push_back(B->fold ? refs[0]->d_qfold : refs[0]->d_qhat);
        nvalid.push_back(0);
        bkt0.push_back(0);
        ++cb;
    }
    B->ncb = cb;
    B->ldA = cb * 64;
    int rc = 0;
    const size_t nrt_max = (size_t)(b->R_pad16 / 64);
    if ((rc = dev_alloc(&B->d_ev, B->ne)) || (rc = dev_alloc(&B->d_gs, ngroups)) ||
        (rc = dev_alloc(&B->d_prop, (size_t)B->ne * B->cstride)) ||
        (sweep && (rc = dev_alloc(&B->d_cur, (size_t)ngroups * B->cstride))) ||
        (rc = dev_alloc(&B->d_quat, (size_t)B->ne * std::max(1, B->nl) * 4)) ||
        (!B->sym && (rc = dev_alloc(&B->d_As, (size_t)b->n_lt * B->ldA))) ||
        (B->sym && (rc = dev_alloc(&B->d_As32, (size_t)b->n_lt * B->ldA))) || (rc = dev_alloc(&B->d_sc, B->ldA)) ||
        (rc = dev_alloc(&B->d_isc, B->ldA)) || (rc = dev_alloc(&B->d_part, nrt_max * 2 * B->ldA)) ||
        (rc = dev_alloc(&B->d_E, B->ne)) || (rc = dev_alloc(&B->d_Ehist, (size_t)std::max(1, hist_cap) * ngroups)) ||
        (rc = dev_alloc(&B->d_shsc, (size_t)B->ne * b->K)) ||
        (!fac && (rc = dev_alloc(&B->d_Ahi, (size_t)B->ldA * b->ldk))) ||
        (!fac && B->prec != TP_PREC_BF16 && (rc = dev_alloc(&B->d_Alo, (size_t)B->ldA * b->ldk))) ||
        (fac && (rc = dev_alloc(&B->d_G, (size_t)(b->lam ? b->n_m : 16 * b->nrb16) * B->ldA))) ||
        (b->lam && (rc = dev_alloc(&B->d_bkt0, bkt0.size()))) ||
        (rc = dev_alloc(&B->d_qtab, qtab.size())) || (rc = dev_alloc(&B->d_nvalid, nvalid.size())) ||
        (!sweep && (rc = dev_alloc(&B->d_raw, (size_t)b->K * b->LMAX * b->LMAX)))) {
        batch_free(B);
        return rc;
    }
    for (int e = 0; e < B->ne; ++e) {
        EvalDesc &d = B->h_ev[e];
        d.coef = B->d_prop + (size_t)e * B->cstride;
        d.src = sweep ? B->d_cur + (size_t)d.group * B->cstride : nullptr;
    }
    B->items.resize(B->ne);
    for (int e = 0; e < B->ne; ++e) batch_items(b, e, B->h_ev[e].K, B->items[e]);
    hipError_t e = hipMemcpy(B->d_ev, B->h_ev.data(), sizeof(EvalDesc) * B->ne, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_qtab, qtab.data(), sizeof(double *) * qtab.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(B->d_nvalid, nvalid.data(), sizeof(int) * nvalid.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && B->d_bkt0)
        e = hipMemcpy(B->d_bkt0, bkt0.data(), sizeof(int) * bkt0.size(), hipMemcpyHostToDevice);
    // As columns that no evaluation writes (tile padding) must read as zero
    if (e == hipSuccess && B->d_As) e = hipMemset(B->d_As, 0, sizeof(double) * (size_t)b->n_lt * B->ldA);
    if (e == hipSuccess && B->d_As32) e = hipMemset(B->d_As32, 0, sizeof(float) * (size_t)b->n_lt * B->ldA);
    if (e == hipSuccess) e = hipEventCreate(&B->ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&B->ev[1]);
    if (e != hipSuccess) {
        batch_free(B);
        return tp_fail(TP_E_HIP, std::string("batch set-up: ") + hipGetErrorString(e));
    }
    if ((rc = batch_upload_work(B, 0)) || (rc = batch_sym_work(B, 0))) {
        batch_free(B);
        return rc;
    }
    *out = B;
    return 0;
}

template <int BM, int BN, int WALK>
static void batch_gemm_launch(const tp_batch *B, hipStream_t st)
{
    const tp_basis *b = B->b;
    const int n_rt = (int)(b->R_pad16 / BM), n_ct = (int)(B->ldA / BN);
    // WALK = 2: one workgroup per column tile walking all row tiles; 1: one per
    // tile; 3: one per tile, row tiles partitioned over the XCDs
    const dim3 grid((unsigned)(WALK == 2 ? n_ct : n_rt * n_ct));
    if (B->prec != TP_PREC_BF16)
        hipLaunchKernelGGL((k_gemm16<BM, BN, 1, 0, true, WALK>), grid, dim3(256), 0, st, b->d_Fhi, b->d_Flo,
                           (int64_t)b->ldk, B->d_Ahi, B->d_Alo, (int64_t)b->ldk, b->ldk, b->F_inv, B->d_isc, (int)b->R,
                           0, n_rt, n_ct, nullptr, (int64_t)0, nullptr, B->d_part, B->ldA, B->d_qtab, B->d_nvalid);
    else
        hipLaunchKernelGGL((k_gemm16<BM, BN, 2, 0, true, WALK>), grid, dim3(256), 0, st, b->d_Fbf, nullptr,
                           (int64_t)b->ldk, B->d_Ahi, nullptr, (int64_t)b->ldk, b->ldk, b->F_inv, B->d_isc, (int)b->R,
                           0, n_rt, n_ct, nullptr, (int64_t)0, nullptr, B->d_part, B->ldA, B->d_qtab, B->d_nvalid);
}

// Contraction candidates: tile (rows x columns) and schedule (1: parallel row
// tiles, 2: row-walking workgroups).
static const int kTiles[kNumTiles][3] = {{128, 64, 1}, {64, 64, 1}, {128, 128, 1}, {64, 128, 1},
                                         {128, 64, 2}, {64, 64, 2}, {128, 128, 2}, {64, 128, 2},
                                         {128, 64, 3}, {64, 64, 3}, {128, 128, 3}, {64, 128, 3}};

static void batch_gemm(const tp_batch *B, hipStream_t st, int cand)
{
    switch (cand) {
    case 0: batch_gemm_launch<128, 64, 1>(B, st); break;
    case 1: batch_gemm_launch<64, 64, 1>(B, st); break;
    case 2: batch_gemm_launch<128, 128, 1>(B, st); break;
    case 3: batch_gemm_launch<64, 128, 1>(B, st); break;
    case 4: batch_gemm_launch<128, 64, 2>(B, st); break;
    case 5: batch_gemm_launch<64, 64, 2>(B, st); break;
    case 6: batch_gemm_launch<128, 128, 2>(B, st); break;
    case 7: batch_gemm_launch<64, 128, 2>(B, st); break;
    case 8: batch_gemm_launch<128, 64, 3>(B, st); break;
    case 9: batch_gemm_launch<64, 64, 3>(B, st); break;
    case 10: batch_gemm_launch<128, 128, 3>(B, st); break;
    default: batch_gemm_launch<64, 128, 3>(B, st); break;
    }
}

// Factored contraction candidates (TP_PREC_FP16X3_T64): k_hgemm_g with 64
// columns x all row blocks, 32 columns x all row blocks, 64 columns x 3 row
// blocks, 16 columns x 4 row blocks per workgroup; then k_c3sep over the
// batch's columns.
constexpr int kNumFac = 4;
static const int kFacTiles[kNumFac][1] = {{64}, {32}, {64}, {16}};

template <int BN, int NRB>
static void batch_hgemm_launch(const tp_batch *B, hipStream_t st)
{
    const tp_basis *b = B->b;
    const unsigned ny = (unsigned)((b->nrb16 + NRB - 1) / NRB);
    hipLaunchKernelGGL((k_hgemm_g<BN, NRB, false, 32, true>), dim3((unsigned)(B->ldA / BN), ny), dim3(256), 0, st,
                       b->d_Hhi, b->d_Hlo, b->ldh16, b->H_inv, nullptr, B->ldA, (int)b->n_lt, b->ldh16, B->d_sc,
                       nullptr, b->L, b->K, b->n_m, B->d_G, B->ldA, B->d_As32);
}

// rows (classes) per workgroup and columns of factored candidate cand
static void fac_shape(const tp_basis *b, int cand, int *rows, int *cols)
{
    const int nrb = b->nrb16;
    const int nr[kNumFac] = {nrb <= 9 ? 9 : 12, nrb <= 10 ? 10 : 12, 3, 4};
    *rows = 16 * nr[cand];
    *cols = kFacTiles[cand][0];
}

static void batch_contract_fac(const tp_batch *B, hipStream_t st, int cand)
{
    const tp_basis *b = B->b;
    if (b->lam) {
        // curved Ewald sphere: G per column (k_glam, FP64), then c3 = Bm . G
        // on the separated grid (the wide instance for the odd-m classes)
        hipLaunchKernelGGL(k_glam_batch, dim3((unsigned)b->n_m, (unsigned)((B->ldA + b->lam_nt - 1) / b->lam_nt)),
                           dim3((unsigned)b->lam_nt), b->lam_lds, st, B->d_As32, B->ldA, b->d_ktk, (int)b->klength,
                           b->d_lamY, b->K, b->d_lamyr, b->d_lamys, b->lam_rows, b->d_lament, b->d_lamcs, B->d_G,
                           B->ldA,
                           B->d_nvalid, B->d_bkt0);
        const int64_t tap = (int64_t)b->n_m * b->jpad;
        const dim3 grid((unsigned)(B->ldA / 16), (unsigned)b->njc);
        if (b->sep_wide)
            hipLaunchKernelGGL((k_c3sep<0, true, true, false, kSepWideQB>), grid, dim3(512),
                               c3sep_lds_bytes(b->sep_kb), st, B->d_G, B->ldA, b->d_m2start, b->n_m2, b->d_TA, tap,
                               b->jpad, b->d_TB, b->sep_kb, b->N, 0, nullptr, B->d_part, B->ldA, nullptr, (int64_t)0,
                               b->d_logtab, B->d_qtab, B->d_nvalid);
        else
            hipLaunchKernelGGL((k_c3sep<0, true, true>), grid, dim3(512), c3sep_lds_bytes(b->sep_kb), st, B->d_G,
                               B->ldA, b->d_m2start, b->n_m2, b->d_TA, tap, b->jpad, b->d_TB, b->sep_kb, b->N, 0,
                               nullptr, B->d_part, B->ldA, nullptr, (int64_t)0, b->d_logtab, B->d_qtab,
                               B->d_nvalid);
        return;
    }
    const int nrb = b->nrb16;
    switch (cand) {
    case 0: if (nrb <= 9) batch_hgemm_launch<64, 9>(B, st); else batch_hgemm_launch<64, 12>(B, st); break;
    case 1: if (nrb <= 10) batch_hgemm_launch<32, 10>(B, st); else batch_hgemm_launch<32, 12>(B, st); break;
    case 2: batch_hgemm_launch<64, 3>(B, st); break;
    default: batch_hgemm_launch<16, 4>(B, st); break;
    }
    const int64_t tap = (int64_t)b->n_m * b->jpad;
    if (B->fold)
        hipLaunchKernelGGL((k_c3sep<0, true, true, true>), dim3((unsigned)(B->ldA / 16), (unsigned)b->njf), dim3(512),
                           c3sep_lds_bytes(b->sep_kb), st, B->d_G, B->ldA, b->d_m2start, b->n_m2, b->d_TA, tap,
                           b->jpad, b->d_TB, b->sep_kb, b->N, 0, nullptr, B->d_part, B->ldA, nullptr, (int64_t)0,
                           b->d_logtab, B->d_qtab, B->d_nvalid);
    else
        hipLaunchKernelGGL((k_c3sep<0, true, true>), dim3((unsigned)(B->ldA / 16), (unsigned)b->njc), dim3(512),
                           c3sep_lds_bytes(b->sep_kb), st, B->d_G, B->ldA, b->d_m2start, b->n_m2, b->d_TA, tap,
                           b->jpad, b->d_TB, b->sep_kb, b->N, 0, nullptr, B->d_part, B->ldA, nullptr, (int64_t)0,
                           b->d_logtab, B->d_qtab, B->d_nvalid);
}

// energy-sum rows the chosen contraction leaves in part
static int batch_part_rows(const tp_batch *B)
{
    if (B->prec == TP_PREC_FP16X3_T64) return B->fold ? B->b->njf : B->b->njc;
    return kTiles[B->gcand][2] == 2 ? 1 : (int)(B->b->R_pad16 / kTiles[B->gcand][0]);
}

// stages of one batched evaluation (coefficients already in place)
static void batch_triple(const tp_batch *B, hipStream_t st)
{
    const tp_basis *b = B->b;
    hipLaunchKernelGGL(k_colscale_batch, dim3((unsigned)B->ne), dim3(256), 0, st, B->d_ev, b->n_c, b->L, B->d_sc,
                       B->d_isc, B->d_shsc, b->K);
    if (B->sym)
        hipLaunchKernelGGL(k_triple_sym_batch, dim3((unsigned)B->nswork), dim3(kSymThreadsBatch), b->sym_lds_b, st, B->d_ev,
                           B->d_swork, B->d_prop, (int64_t)B->cstride, b->n_c, b->d_ew, b->d_eidx, b->d_gstart,
                           b->d_perms, b->d_gorder,
                           b->d_symtab, b->d_symoff, b->Lw, B->d_As32, B->ldA);
    else
        hipLaunchKernelGGL(k_triple16_batch, dim3((unsigned)B->nwork), dim3(512), b->t16_lds, st, B->d_ev, B->d_work,
                           B->d_shsc, b->K, b->n_c, b->d_ew8, b->d_eidx8, b->d_gstart8, b->d_ltinfo, b->Lw, b->EmaxLt8,
                           b->Prows, B->d_As, B->ldA);
}

static void batch_split(const tp_batch *B, hipStream_t st)
{
    const tp_basis *b = B->b;
    const dim3 gs((unsigned)B->ncb, (unsigned)(b->ldk / 64));
    if (B->sym && B->prec != TP_PREC_BF16)
        hipLaunchKernelGGL((k_split_batch<1, float>), gs, dim3(256), 0, st, B->d_As32, B->ldA, (int)b->n_lt, b->ldk,
                           B->d_sc, B->d_Ahi, B->d_Alo);
    else if (B->sym)
        hipLaunchKernelGGL((k_split_batch<2, float>), gs, dim3(256), 0, st, B->d_As32, B->ldA, (int)b->n_lt, b->ldk,
                           B->d_sc, B->d_Ahi, nullptr);
    else if (B->prec != TP_PREC_BF16)
        hipLaunchKernelGGL(k_split_batch<1>, gs, dim3(256), 0, st, B->d_As, B->ldA, (int)b->n_lt, b->ldk, B->d_sc,
                           B->d_Ahi, B->d_Alo);
    else
        hipLaunchKernelGGL(k_split_batch<2>, gs, dim3(256), 0, st, B->d_As, B->ldA, (int)b->n_lt, b->ldk, B->d_sc,
                           B->d_Ahi, nullptr);
}

static void batch_final(const tp_batch *B, hipStream_t st)
{
    hipLaunchKernelGGL(k_final_batch, dim3((unsigned)B->ne), dim3(256), 0, st, B->d_ev, B->d_part,
                       batch_part_rows(B), B->ldA, B->d_E);
}

static int batch_enqueue_eval(const tp_batch *B, hipStream_t st)
{
    batch_triple(B, st);
    if (B->prec == TP_PREC_FP16X3_T64) {
        batch_contract_fac(B, st, B->gcand);
    } else {
        batch_split(B, st);
        batch_gemm(B, st, B->gcand);
    }
    batch_final(B, st);
    TP_HIP(hipGetLastError());
    return 0;
}

static void batch_learned_quats(const tp_batch *B, hipStream_t st);   // learned slots (defined with the MLP code)

// One sweep: proposals (random quaternions; learned ones in the last nlearn
// slots), rotation, evaluation, move log, decision.
static int batch_enqueue_sweep(const tp_batch *B, hipStream_t st, int probe, double decay, double factor, double tfac)
{
    const tp_basis *b = B->b;
    const int nq = B->ne * B->nl;
    hipLaunchKernelGGL(k_quat_random, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, B->d_ev, B->ne, B->nl,
                       B->d_gs, probe, B->d_quat);
    if (!probe && B->mlp && B->nlearn > 0) batch_learned_quats(B, st);
    hipLaunchKernelGGL(k_rotate_batch, dim3((unsigned)B->nl, (unsigned)B->ne), dim3(256), 0, st, B->d_ev, B->d_quat,
                       B->nl, b->n_c, B->d_gs, B->family, probe);
    if (int rc = batch_enqueue_eval(B, st)) return rc;
    if (!probe && B->log_cap > 0)
        hipLaunchKernelGGL(k_log_moves, dim3(1), dim3(1024), 0, st, B->d_ev, B->ne, B->nl, B->d_gs, B->d_E, B->d_quat,
                           b->n_c, B->d_Xlog, B->d_Tlog, B->d_logn, B->log_cap);
    hipLaunchKernelGGL(k_select_batch, dim3((unsigned)B->ngroups), dim3(256), 0, st, B->d_gs, B->d_ev, B->d_E,
                       B->d_cur, B->cstride, b->n_c, probe, decay, factor, tfac, B->d_Ehist, B->hist_cap,
                       B->ngroups);
    TP_HIP(hipGetLastError());
    return 0;
}

static int batch_capture(tp_batch *B, double decay, double factor)
{
    hipStream_t st = B->ctx->stream;
    if (B->exec) {
        TP_HIP(hipGraphExecDestroy(B->exec));
        B->exec = nullptr;
    }
    if (B->graph) {
        TP_HIP(hipGraphDestroy(B->graph));
        B->graph = nullptr;
    }
    TP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = B->sweep ? batch_enqueue_sweep(B, st, 0, decay, factor, 0.0) : batch_enqueue_eval(B, st);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return tp_fail(TP_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    B->graph = g;
    TP_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
    B->g_decay = decay;
    B->g_factor = factor;
    return 0;
}

// Time `reps` launches of a stage on the context stream (ms per launch).
template <class F>
static int batch_time(tp_batch *B, int reps, F &&f, double *ms)
{
    hipStream_t st = B->ctx->stream;
    f(st);
    TP_HIP(hipGetLastError());
    TP_HIP(hipEventRecord(B->ev[0], st));
    for (int i = 0; i < reps; ++i) f(st);
    TP_HIP(hipEventRecord(B->ev[1], st));
    TP_HIP(hipEventSynchronize(B->ev[1]));
    TP_HIP(hipGetLastError());
    float t = 0.f;
    TP_HIP(hipEventElapsedTime(&t, B->ev[0], B->ev[1]));
    *ms = t / reps;
    return 0;
}

// Stage sizes of the current coefficients: choose the triple-product work
// order and the contraction tile by timing every candidate on this batch.
static int batch_autotune(tp_batch *B, int reps)
{
    hipStream_t st = B->ctx->stream;
    double best = 1e30;
    int best_order = 0;
    // work orders: of the sorted l-triple form, or (bases from a caller's weights) of the coupling form
    const bool sym = B->sym;
    for (int order = 0; order < 2; ++order) {
        B->tune_ms[order] = 1e30;
        if (int rc = sym ? batch_sym_work(B, order) : batch_upload_work(B, order)) return rc;
        double ms = 0;
        if (int rc = batch_time(B, reps, [&](hipStream_t s) { batch_triple(B, s); }, &ms)) return rc;
        B->tune_ms[order] = ms;
        if (ms < best) { best = ms; best_order = order; }
    }
    if (int rc = sym ? batch_sym_work(B, best_order) : batch_upload_work(B, best_order)) return rc;
    batch_triple(B, st);
    best = 1e30;
    int best_t = 0;
    if (B->prec == TP_PREC_FP16X3_T64) {
        for (int t = 0; t < kNumTiles; ++t) {
            B->tune_ms[2 + t] = 1e30;
            if (t >= kNumFac) continue;
            double ms = 0;
            if (int rc = batch_time(B, reps, [&](hipStream_t s) { batch_contract_fac(B, s, t); }, &ms)) return rc;
            B->tune_ms[2 + t] = ms;
            if (ms < best) { best = ms; best_t = t; }
        }
        B->gcand = best_t;
        TP_HIP(hipStreamSynchronize(st));
        if (B->exec) return batch_capture(B, B->g_decay, B->g_factor);
        return 0;
    }
    batch_split(B, st);
    const int nct64 = (int)(B->ldA / 64);
    for (int t = 0; t < kNumTiles; ++t) {
        double ms = 0;
        if (kTiles[t][2] == 3) {   // the XCD partition needs n_rt and the grid divisible by 8
            const int nrt = (int)(B->b->R_pad16 / kTiles[t][0]), nct = (int)(B->ldA / kTiles[t][1]);
            if (nrt % 8 != 0 || (nrt * nct) % 8 != 0 || nct64 < 1) {
                B->tune_ms[2 + t] = 1e30;
                continue;
            }
        }
        if (int rc = batch_time(B, reps, [&](hipStream_t s) { batch_gemm(B, s, t); }, &ms)) return rc;
        B->tune_ms[2 + t] = ms;
        if (ms < best) { best = ms; best_t = t; }
    }
    B->gcand = best_t;
    TP_HIP(hipStreamSynchronize(st));
    if (B->exec) return batch_capture(B, B->g_decay, B->g_factor);
    return 0;
}

// ---------------------------------------------------------------------------
// learned proposer (MLP, tp_mlp.hpp)
// ---------------------------------------------------------------------------
struct tp_mlp {
    tp_ctx *ctx = nullptr;
    float *d_p = nullptr, *d_m = nullptr, *d_v = nullptr, *d_g = nullptr;   // [kMlpNParam]
    uint16_t *d_W[4] = {nullptr, nullptr, nullptr, nullptr}, *d_Wt[4] = {nullptr, nullptr, nullptr, nullptr};
    int *d_t = nullptr;            // Adam steps taken
    // activations / gradients for up to Bcap rows
    int Bcap = 0;
    uint16_t *X = nullptr, *Xt = nullptr, *H[3] = {nullptr, nullptr, nullptr}, *Ht[3] = {nullptr, nullptr, nullptr};
    uint16_t *dY = nullptr, *dYt = nullptr, *dH[3] = {nullptr, nullptr, nullptr}, *dHt[3] = {nullptr, nullptr, nullptr};
    float *Y = nullptr, *T = nullptr, *lpart = nullptr, *trace = nullptr;
    int trace_cap = 0;
    // host-supplied training data
    uint16_t *d_Xdata = nullptr;
    float *d_Tdata = nullptr;
    int *d_count = nullptr;
    int ndata_cap = 0;
    // captured training step
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void *g_src = nullptr, *g_tgt = nullptr, *g_cnt = nullptr;   // the buffers the graph was captured on
    int g_B = 0;
    float g_lr = 0.f;
    uint64_t g_seed = 0;
};

static int mlp_bias_off(int L) { return kMlpOffW[L] + kMlpRows[L] * kMlpCols[L]; }

static void mlp_free_acts(tp_mlp *m)
{
    void *ptrs[] = {m->X, m->Xt, m->H[0], m->H[1], m->H[2], m->Ht[0], m->Ht[1], m->Ht[2], m->dY, m->dYt,
                    m->dH[0], m->dH[1], m->dH[2], m->dHt[0], m->dHt[1], m->dHt[2], m->Y, m->T, m->lpart};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    m->X = m->Xt = m->dY = m->dYt = nullptr;
    for (int i = 0; i < 3; ++i) m->H[i] = m->Ht[i] = m->dH[i] = m->dHt[i] = nullptr;
    m->Y = m->T = m->lpart = nullptr;
    m->Bcap = 0;
}

static int mlp_ensure(tp_mlp *m, int B)
{
    if (B <= m->Bcap) return 0;
    mlp_free_acts(m);
    if (m->exec) { (void)hipGraphExecDestroy(m->exec); m->exec = nullptr; }
    if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
    int rc = 0;
    const size_t b = (size_t)B;
    if ((rc = dev_alloc(&m->X, b * kMlpIn)) || (rc = dev_alloc(&m->Xt, b * kMlpIn)) ||
        (rc = dev_alloc(&m->dY, b * kMlpOut)) || (rc = dev_alloc(&m->dYt, b * kMlpOut)) ||
        (rc = dev_alloc(&m->Y, b * kMlpOut)) || (rc = dev_alloc(&m->T, b * kMlpQ)) || (rc = dev_alloc(&m->lpart, b)))
        return rc;
    for (int i = 0; i < 3; ++i)
        if ((rc = dev_alloc(&m->H[i], b * kMlpHid)) || (rc = dev_alloc(&m->Ht[i], b * kMlpHid)) ||
            (rc = dev_alloc(&m->dH[i], b * kMlpHid)) || (rc = dev_alloc(&m->dHt[i], b * kMlpHid)))
            return rc;
    TP_HIP(hipMemset(m->X, 0, sizeof(uint16_t) * b * kMlpIn));
    m->Bcap = B;
    return 0;
}

static MlpEpi epi_none()
{
    MlpEpi e;
    std::memset(&e, 0, sizeof(e));
    e.scale = 1.f;
    return e;
}

static void mlp_gemm(hipStream_t st, const uint16_t *A, int64_t lda, const uint16_t *B, int64_t ldb, int M, int N, int K,
                     const MlpEpi &e)
{
    hipLaunchKernelGGL(k_mlp_gemm, dim3((unsigned)((M / 64) * (N / 64))), dim3(256), 0, st, A, lda, B, ldb, M, N, K, e);
}

// Forward pass of M rows of X (BF16 [M][192]); with `train` the transposed
// activations are kept for the backward pass.  Hidden layers in H (BF16),
// outputs in Y (FP32 [M][64]).
static void mlp_forward(const tp_mlp *m, hipStream_t st, const uint16_t *X, int M, uint16_t *const H[3],
                        uint16_t *const Ht[3], float *Y)
{
    const uint16_t *in = X;
    int64_t ldin = kMlpIn;
    for (int L = 0; L < 3; ++L) {
        MlpEpi e = epi_none();
        e.bias = m->d_p + mlp_bias_off(L);
        e.relu = 1;
        e.outb = H[L];
        e.ldo = kMlpHid;
        if (Ht) { e.outt = Ht[L]; e.ldt = M; }
        mlp_gemm(st, in, ldin, m->d_W[L], kMlpCols[L], M, kMlpHid, kMlpCols[L], e);
        in = H[L];
        ldin = kMlpHid;
    }
    MlpEpi e = epi_none();
    e.bias = m->d_p + mlp_bias_off(3);
    e.outf = Y;
    e.ldf = kMlpOut;
    mlp_gemm(st, in, ldin, m->d_W[3], kMlpHid, M, kMlpOut, kMlpHid, e);
}

// Loss and gradient of B rows (X, Xt, T already in place) into d_g.
static void mlp_backward(const tp_mlp *m, hipStream_t st, int B)
{
    hipLaunchKernelGGL(k_mlp_loss, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, st, m->Y, m->T, B, m->dY, m->dYt,
                       m->lpart);
    const uint16_t *dcur = m->dY, *dcurt = m->dYt;
    int ncur = kMlpOut;
    for (int L = 3; L >= 0; --L) {
        // dW_L [rows][cols] = dcur^T . act_L^T   (act_0 = X)
        const uint16_t *actt = L == 0 ? m->Xt : m->Ht[L - 1];
        MlpEpi e = epi_none();
        e.outf = m->d_g + kMlpOffW[L];
        e.ldf = kMlpCols[L];
        mlp_gemm(st, dcurt, B, actt, B, kMlpRows[L], kMlpCols[L], B, e);
        hipLaunchKernelGGL(k_mlp_rowsum, dim3((unsigned)((kMlpRows[L] + 3) / 4)), dim3(256), 0, st, dcurt, (int64_t)B,
                           B, m->d_g + mlp_bias_off(L), kMlpRows[L]);
        if (L == 0) break;
        // dH_{L-1} = (dcur . W_L) * [H_{L-1} > 0]
        MlpEpi e2 = epi_none();
        e2.mask = m->H[L - 1];
        e2.ldm = kMlpHid;
        e2.outb = m->dH[L - 1];
        e2.ldo = kMlpHid;
        e2.outt = m->dHt[L - 1];
        e2.ldt = B;
        mlp_gemm(st, dcur, ncur, m->d_Wt[L], kMlpRows[L], B, kMlpCols[L], kMlpRows[L], e2);
        dcur = m->dH[L - 1];
        dcurt = m->dHt[L - 1];
        ncur = kMlpHid;
    }
}

static MlpWeights mlp_weights(const tp_mlp *m)
{
    MlpWeights w;
    for (int L = 0; L < 4; ++L) { w.W[L] = m->d_W[L]; w.Wt[L] = m->d_Wt[L]; }
    return w;
}

// Refresh the BF16 copies from the master parameters (an Adam step with lr 0
// and zero moments would do the same; this is a plain conversion).
static int mlp_upload(tp_mlp *m, const std::vector<float> &p)
{
    TP_HIP(hipMemcpy(m->d_p, p.data(), sizeof(float) * kMlpNParam, hipMemcpyHostToDevice));
    for (int L = 0; L < 4; ++L) {
        const int R = kMlpRows[L], Cc = kMlpCols[L];
        std::vector<uint16_t> w((size_t)R * Cc), wt((size_t)R * Cc);
        for (int r = 0; r < R; ++r)
            for (int c = 0; c < Cc; ++c) {
                const uint16_t h = host_bf16(p[(size_t)kMlpOffW[L] + (size_t)r * Cc + c]);
                w[(size_t)r * Cc + c] = h;
                wt[(size_t)c * R + r] = h;
            }
        TP_HIP(hipMemcpy(m->d_W[L], w.data(), sizeof(uint16_t) * w.size(), hipMemcpyHostToDevice));
        TP_HIP(hipMemcpy(m->d_Wt[L], wt.data(), sizeof(uint16_t) * wt.size(), hipMemcpyHostToDevice));
    }
    TP_HIP(hipMemset(m->d_m, 0, sizeof(float) * kMlpNParam));
    TP_HIP(hipMemset(m->d_v, 0, sizeof(float) * kMlpNParam));
    TP_HIP(hipMemset(m->d_t, 0, sizeof(int)));
    return 0;
}

static void mlp_free(tp_mlp *m)
{
    if (!m) return;
    if (m->ctx) (void)hipSetDevice(m->ctx->device);
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    mlp_free_acts(m);
    void *ptrs[] = {m->d_p, m->d_m, m->d_v, m->d_g, m->d_t, m->trace, m->d_Xdata, m->d_Tdata, m->d_count};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (int L = 0; L < 4; ++L) {
        if (m->d_W[L]) (void)hipFree(m->d_W[L]);
        if (m->d_Wt[L]) (void)hipFree(m->d_Wt[L]);
    }
    delete m;
}

__global__ void k_mlp_losssum(const float *lpart, int B, const int *tstep, float *trace, int cap)
{
    if (threadIdx.x != 0) return;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += lpart[b];
    const int t = tstep[0];
    if (t < cap) trace[t] = s;
}

// One training step on data rows (Xd, Td, count): gather a minibatch, forward,
// backward, Adam, loss to trace[t], t += 1.
static void mlp_train_step(tp_mlp *m, hipStream_t st, const uint16_t *Xd, const float *Td, const int *count, int B,
                           float lr, uint64_t seed)
{
    hipLaunchKernelGGL(k_mlp_gather, dim3((unsigned)B), dim3(256), 0, st, Xd, Td, count, seed, m->d_t, B, m->X, m->Xt,
                       m->T);
    mlp_forward(m, st, m->X, B, m->H, m->Ht, m->Y);
    mlp_backward(m, st, B);
    hipLaunchKernelGGL(k_mlp_losssum, dim3(1), dim3(64), 0, st, m->lpart, B, m->d_t, m->trace, m->trace_cap);
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)((kMlpNParam + 255) / 256)), dim3(256), 0, st, m->d_p, m->d_m,
                       m->d_v, m->d_g, m->d_t, lr, 0.9f, 0.999f, 1e-8f, mlp_weights(m));
    hipLaunchKernelGGL(k_mlp_tick, dim3(1), dim3(1), 0, st, m->d_t);
}

static int mlp_train(tp_mlp *m, const uint16_t *Xd, const float *Td, const int *count, int steps, int B, float lr,
                     uint64_t seed, float *loss_trace)
{
    hipStream_t st = m->ctx->stream;
    if (int rc = mlp_ensure(m, B)) return rc;
    int t0 = 0;
    TP_HIP(hipMemcpy(&t0, m->d_t, sizeof(int), hipMemcpyDeviceToHost));
    if (m->trace_cap < t0 + steps) {
        if (m->trace) TP_HIP(hipFree(m->trace));
        m->trace = nullptr;
        m->trace_cap = 0;
        if (int rc = dev_alloc(&m->trace, (size_t)(t0 + steps) * 2)) return rc;
        m->trace_cap = (t0 + steps) * 2;
        if (m->exec) { TP_HIP(hipGraphExecDestroy(m->exec)); m->exec = nullptr; }
        if (m->graph) { TP_HIP(hipGraphDestroy(m->graph)); m->graph = nullptr; }
    }
    if (!m->exec || m->g_src != (const void *)Xd || m->g_tgt != (const void *)Td || m->g_cnt != (const void *)count ||
        m->g_B != B || m->g_lr != lr || m->g_seed != seed) {
        if (m->exec) { TP_HIP(hipGraphExecDestroy(m->exec)); m->exec = nullptr; }
        if (m->graph) { TP_HIP(hipGraphDestroy(m->graph)); m->graph = nullptr; }
        TP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        mlp_train_step(m, st, Xd, Td, count, B, lr, seed);
        hipGraph_t g = nullptr;
        TP_HIP(hipStreamEndCapture(st, &g));
        m->graph = g;
        TP_HIP(hipGraphInstantiate(&m->exec, g, nullptr, nullptr, 0));
        m->g_src = Xd;
        m->g_tgt = Td;
        m->g_cnt = count;
        m->g_B = B;
        m->g_lr = lr;
        m->g_seed = seed;
    }
    for (int i = 0; i < steps; ++i) TP_HIP(hipGraphLaunch(m->exec, st));
    if (loss_trace)
        TP_HIP(hipMemcpyAsync(loss_trace, m->trace + t0, sizeof(float) * steps, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    return 0;
}

static uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void batch_learned_quats(const tp_batch *B, hipStream_t st)
{
    const tp_basis *b = B->b;
    hipLaunchKernelGGL(k_mlp_infer_features, dim3((unsigned)B->nlearn, (unsigned)B->ngroups), dim3(64), 0, st,
                       B->d_gs, B->d_cur, B->cstride, b->n_c, B->nlearn, B->d_Xi);
    uint16_t *const H[3] = {B->d_Hi[0], B->d_Hi[1], B->d_Hi[2]};
    mlp_forward(B->mlp, st, B->d_Xi, B->Mi, H, nullptr, B->d_Yi);
    const int nq = B->ne * B->nl;
    hipLaunchKernelGGL(k_mlp_to_quat, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, B->d_ev, B->ne, B->nl,
                       B->d_gs, B->d_Yi, B->noise, B->d_quat);
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
    if (device >= kMaxDevices) return tp_fail(TP_E_UNSUPPORTED, "device ordinal above 63");
    TP_HIP(hipSetDevice(device));
    void *scr = nullptr;
    {
        // once per device, with that device current: its own copy of the array
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[device]) {
            e = hipGetSymbolAddress(&g_scratch_ptr[device], HIP_SYMBOL(scratch));
            if (e != hipSuccess) {
                g_scratch_ptr[device] = nullptr;
                return tp_fail(TP_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            }
        }
        scr = g_scratch_ptr[device];
    }
    tp_ctx *c = new tp_ctx();
    c->device = device;
    c->scratch = scr;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return tp_fail(TP_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < 5; ++i) {
        e = hipEventCreate(&c->ev[i]);
        if (e != hipSuccess) {
            tp_finalize(c);
            return tp_fail(TP_E_HIP, "hipEventCreate failed");
        }
    }
    e = hipHostMalloc((void **)&c->h_pin, 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        c->h_pin = nullptr;
        tp_finalize(c);
        return tp_fail(TP_E_HIP, "hipHostMalloc failed");
    }
    // the energy of a synchronous evaluation is stored by k_final directly in
    // host memory: no copy (a blit kernel for small copies) after the evaluation
    e = hipHostMalloc((void **)&c->h_res, 64, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void **)&c->d_res, c->h_res, 0);
    if (e != hipSuccess) {
        tp_finalize(c);
        return tp_fail(TP_E_HIP, std::string("mapped result buffer: ") + hipGetErrorString(e));
    }
    *out = c;
    return 0;
}

void tp_finalize(tp_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < 5; ++i)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->h_pin) (void)hipHostFree(ctx->h_pin);
    if (ctx->h_res) (void)hipHostFree(ctx->h_res);
    batch_free(ctx->eb);
    ph_cache_free(ctx);
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
    for (int i = 0; i < n && i < 4; ++i) ms_out[i] = ctx->last_ms[i];
    return 0;
}

void tp_basis_destroy(tp_basis *b)
{
    if (!b) return;
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    void *ptrs[] = {b->d_H,   b->d_Bm,  b->d_ew,   b->d_eidx,    b->d_gstart, b->d_ltinfo, b->d_chunks,
                    b->d_pairs, b->d_sean0, b->d_coff, b->d_As, b->d_G, b->d_part, b->d_out, b->d_raw, b->d_coef, b->d_c3, b->d_work, b->d_Fhi, b->d_Flo,
                    b->d_Fbf, b->d_Ahi, b->d_Alo, b->d_Ainv, b->d_shsc, b->d_ew8, b->d_eidx8, b->d_gstart8,
                    b->d_m2start, b->d_TA, b->d_TB, b->d_logtab, b->d_HT4, b->d_gGp, b->d_gA, b->d_gpart,
                    b->d_grad, b->d_lstart, b->d_llist, b->d_bwd_order, b->d_c2coef, b->d_c2raw, b->d_c2,
                    b->d_lamY, b->d_ktk, b->d_lament, b->d_lamcs, b->d_lamyr, b->d_lamys,
                    b->d_lambent, b->d_lambcs, b->d_lambyr, b->d_lambys, b->d_gG, b->d_sym_work, b->d_perms, b->d_gorder,
                    b->d_symtab, b->d_symoff, b->d_Hhi, b->d_Hlo, b->d_nrm2};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
}

// Weights supplied by the caller (tp_basis_from_host): (l1,l2,l3,m1,m2) -> w.
static uint64_t wkey(int l1, int l2, int l3, int m1, int m2)
{
    return ((((uint64_t)l1 * 64 + l2) * 64 + l3) * 128 + (uint64_t)(m1 + 64)) * 128 + (uint64_t)(m2 + 64);
}

static int basis_build(tp_ctx *ctx, int L, int LMAX, int N, int K, double lambda, double dq, int for_intensity,
                       const std::unordered_map<uint64_t, double> *hw, int64_t expect_len, tp_basis **out);

int tp_basis_create(tp_ctx *ctx, int L, int LMAX, int N, int K, double lambda, double dq, int for_intensity,
                    tp_basis **out)
{
    return basis_build(ctx, L, LMAX, N, K, lambda, dq, for_intensity, nullptr, -1, out);
}

int tp_basis_from_host(tp_ctx *ctx, const double *wignerlist, const int64_t *indices, int64_t basislen,
                       const int64_t *pacombos, int64_t ncombos, int N, int L, int LMAX, int K, double lambda,
                       double dq, tp_basis **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (!wignerlist || !indices || basislen < 1) return tp_fail(TP_E_ARG, "null or empty basis arrays");
    std::unordered_map<uint64_t, double> hw;
    hw.reserve((size_t)basislen * 2);
    for (int64_t j = 0; j < basislen; ++j) {
        const int64_t *t = indices + 9 * j;   // [sean1, sean2, sean3, l1, m1, l2, m2, l3, -m3]
        const int l1 = (int)t[3], m1 = (int)t[4], l2 = (int)t[5], m2 = (int)t[6], l3 = (int)t[7], mu3 = (int)t[8];
        if (l1 < 0 || l2 < 0 || l3 < 0 || l1 > L || l2 > L || l3 > L || std::abs(m1) > l1 || std::abs(m2) > l2 ||
            m1 + m2 + mu3 != 0 || t[0] != tph::seanindex(l1, m1, LMAX) || t[1] != tph::seanindex(l2, m2, LMAX) ||
            t[2] != tph::seanindex(l3, mu3, LMAX))
            return tp_fail(TP_E_ARG, "indices column " + std::to_string(j) + " is not a basis entry");
        hw[wkey(l1, l2, l3, m1, m2)] = wignerlist[j];
    }
    if (pacombos && ncombos > 0) {
        // PAcombos columns [k1,k2,k3,kt,l1,l2,l3,jstart,mcombos] cover every (k-triple, l-triple)
        const int64_t kl = (int64_t)K * (K + 1) * (K + 2) / 6;
        if (ncombos % kl != 0) return tp_fail(TP_E_ARG, "PAcombos length is not a multiple of klength");
    }
    return basis_build(ctx, L, LMAX, N, K, lambda, dq, 1, &hw, basislen, out);
}

static int basis_build(tp_ctx *ctx, int L, int LMAX, int N, int K, double lambda, double dq, int for_intensity,
                       const std::unordered_map<uint64_t, double> *hw, int64_t expect_len, tp_basis **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (int rc = check_ctx(ctx)) return rc;
    if (L < 0 || L > 22 || LMAX < L + 1 || N < 1 || N > 256 || K < 1 || K > 255)
        return tp_fail(TP_E_ARG, "basis parameters out of range (0<=L<=22, LMAX>L, 1<=N<=256, 1<=K<=255)");
    // shell angles theta_k = acos(k lambda dq / (4 pi)) (correlations.jl:94) must exist for k <= K
    if (!(std::fabs(K * lambda * dq / (4 * tph::kPi)) <= 1.0))
        return tp_fail(TP_E_ARG, "shell angles acos(k lambda dq / (4 pi)) undefined: need |K lambda dq| <= 4 pi");
    if (!for_intensity)
        return tp_fail(TP_E_UNSUPPORTED, "for_intensity = false (odd l) is not implemented in this version");
    tp_basis *b = new tp_basis();
    b->ctx = ctx;
    b->L = L; b->LMAX = LMAX; b->N = N; b->K = K; b->lambda = lambda; b->dq = dq;
    b->for_intensity = 1;
    b->lstep = 2;
    b->lam = lambda * dq != 0.0;
    // classes of (m2, mu3): even m only on the flat sphere, where sphPlm(l, m, pi/2) = 0 for odd l + m
    const int mstep = b->lam ? 1 : 2;
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

    // folded (m2, mu3) classes: cos(m2 a + mu3 b) is the same for
    // (m2,mu3) and (-m2,-mu3); representative has m2 > 0, or m2 == 0 and mu3 >= 0
    const int W = 2 * L + 1;
    std::vector<int> cls((size_t)W * W, -1);
    std::vector<int> cls_m2, cls_mu3;
    for (int m2 = 0; m2 <= L; m2 += mstep)
        for (int mu3 = -L; mu3 <= L; mu3 += mstep) {
            if (m2 == 0 && mu3 < 0) continue;
            if (std::abs(m2 + mu3) > L) continue;
            const int c = (int)cls_m2.size();
            cls_m2.push_back(m2);
            cls_mu3.push_back(mu3);
            cls[(size_t)(m2 + L) * W + (mu3 + L)] = c;
            cls[(size_t)(-m2 + L) * W + (-mu3 + L)] = c;
        }
    b->n_m = (int)cls_m2.size();

    // ---- enumeration (correlations.jl:92,107-129): l1, l2, l3, m1, m2, m3 with
    // |3j(l1 m1 l2 m2 l3 -m3)| > 100 eps; stored per (lt, mu3) group sorted by m1.
    const double thr = 100.0 * 2.220446049250313e-16;
    const double th = std::acos(0.0);   // qlist at lambda = 0 (correlations.jl:94)
    std::vector<double> Ytab((size_t)(L + 1) * W);
    for (int l = 0; l <= L; ++l)
        for (int m = -l; m <= l; ++m) Ytab[(size_t)l * W + m + L] = tph::sphPlm(l, m, th);
    auto Y = [&](int l, int m) { return Ytab[(size_t)l * W + m + L]; };
    struct Ent { double w; int i1, i2; };
    std::vector<double> ew;
    std::vector<uint32_t> eidx;
    std::vector<int> gstart, lt_entries, lt_goff;
    std::vector<int4> ltinfo;
    std::vector<std::vector<double>> Hcols;   // per lt: folded class -> H value
    int64_t basislen = 0;
    for (int l1 = 0; l1 <= L; l1 += b->lstep)
        for (int l2 = 0; l2 <= L; l2 += b->lstep)
            for (int l3 = 0; l3 <= L; l3 += b->lstep) {
                if (l3 < std::abs(l1 - l2) || l3 > l1 + l2) continue;
                std::vector<std::vector<Ent>> groups(2 * l3 + 1);
                std::vector<double> hcol(b->n_m, 0.0);
                int cnt = 0;
                for (int m1 = -l1; m1 <= l1; ++m1)
                    for (int m2 = -l2; m2 <= l2; ++m2) {
                        const int mu3 = -(m1 + m2);
                        if (mu3 < -l3 || mu3 > l3) continue;
                        double w;
                        if (hw) {   // the caller's weights, on the reference's own enumeration
                            const auto it = hw->find(wkey(l1, l2, l3, m1, m2));
                            if (it == hw->end()) continue;
                            w = it->second;
                        } else {
                            w = tph::wigner3j(l1, m1, l2, m2, l3, mu3);
                            if (!(std::fabs(w) > thr)) continue;
                        }
                        groups[mu3 + l3].push_back({w, b->coff[l1] + m1 + l1, b->coff[l2] + m2 + l2});
                        // P at theta = pi/2 vanishes for odd m (Y(l,m,pi/2) = 0 when l+m is odd)
                        if (!b->lam && !(m1 & 1) && !(m2 & 1))
                            hcol[cls[(size_t)(m2 + L) * W + (mu3 + L)]] += w * Y(l1, m1) * Y(l2, m2) * Y(l3, mu3);
                        ++cnt;
                    }
                if (cnt == 0) continue;
                lt_goff.push_back((int)gstart.size());
                ltinfo.push_back(make_int4(b->coff[l3], (int)gstart.size(), l3, l1 | (l2 << 8)));
                for (auto &gv : groups) {
                    gstart.push_back((int)ew.size());
                    for (auto &e : gv) {
                        ew.push_back(e.w);
                        eidx.push_back((uint32_t)e.i1 | ((uint32_t)e.i2 << 16));
                    }
                }
                lt_entries.push_back(cnt);
                Hcols.push_back(std::move(hcol));
                basislen += cnt;
            }
    gstart.push_back((int)ew.size());
    const int nlt = (int)lt_entries.size();
    if (hw && basislen != expect_len) {
        delete b;
        return tp_fail(TP_E_ARG, "host basis has entries outside the enumeration of (L, LMAX)");
    }
    // curved Ewald sphere: per-class entries over the l-triples, the shells'
    // sphPlm rows and the k-triple decoding of k_glam
    struct LamEnt {   // host form of an entry, sorted before packing into LamRec
        double hw;
        int lt, y1, y2, y3, fl;
    };
    std::vector<LamEnt> lament;
    std::vector<LamRec> lamrec;
    std::vector<int> lamyr, lamys, lambcs, lambyr, lambys;
    std::vector<LamRec> lambrec;
    std::vector<int> lamcs, ktk;
    std::vector<double> lamY;
    if (b->lam) {
        b->lam_yrow = (L + 1) * W;
        lamY.assign((size_t)K * b->lam_yrow, 0.0);
        for (int k = 1; k <= K; ++k) {
            const double thk = std::acos(k * lambda * dq / (4 * tph::kPi));
            for (int l = 0; l <= L; ++l)
                for (int m = -l; m <= l; ++m) lamY[(size_t)(l * W + m + L) * K + (k - 1)] = tph::sphPlm(l, m, thk);
        }
        for (int c = 0; c < b->n_m; ++c) {
            const size_t first = lament.size();
            lamcs.push_back((int)first);
            const int m2 = cls_m2[c], mu3 = cls_mu3[c], m1 = -(m2 + mu3);
            for (int t = 0; t < nlt; ++t) {
                const int l3 = ltinfo[t].z, l1 = ltinfo[t].w & 255, l2 = ltinfo[t].w >> 8;
                if (std::abs(m1) > l1 || std::abs(m2) > l2 || std::abs(mu3) > l3) continue;
                double w;
                if (hw) {
                    const auto it = hw->find(wkey(l1, l2, l3, m1, m2));
                    if (it == hw->end()) continue;
                    w = it->second;
                } else {
                    w = tph::wigner3j(l1, m1, l2, m2, l3, mu3);
                    if (!(std::fabs(w) > thr)) continue;
                }
                // (-m1, -m2, -mu3) gives the same term: 3j and the sphPlm signs cancel for even l1 + l2 + l3
                const double mult = (m2 == 0 && mu3 == 0) ? 1.0 : 2.0;
                lament.push_back(LamEnt{mult * w, t, (l1 * W + m1 + L) * K, (l2 * W + m2 + L) * K,
                                        (l3 * W + mu3 + L) * K, 0});
            }
            // (l3, l2, l1) order within the class (the rows y3, y2, y1 grow with l)
            std::sort(lament.begin() + first, lament.end(), [](const LamEnt &x, const LamEnt &y) {
                return x.y3 != y.y3 ? x.y3 < y.y3 : (x.y2 != y.y2 ? x.y2 < y.y2 : x.y1 < y.y1);
            });
            for (size_t i = first; i < lament.size(); ++i) {
                const bool n3 = i == first || lament[i].y3 != lament[i - 1].y3;
                const bool n2 = n3 || lament[i].y2 != lament[i - 1].y2;
                lament[i].fl = (n3 ? 1 : 0) | (n2 ? 2 : 0);
            }
        }
        lamcs.push_back((int)lament.size());
        // packed, forward: class c stages the rows (l, m1), (l, m2), (l, mu3)
        // for l = 0, lstep, ..., L (nr rows each; a row with |m| > l is zero)
        const int nr = L / b->lstep + 1;
        int max_ent = 0;
        for (int c = 0; c < b->n_m; ++c) {
            max_ent = std::max(max_ent, lamcs[c + 1] - lamcs[c]);
            lamys.push_back((int)lamyr.size());
            const int ms[3] = {-(cls_m2[c] + cls_mu3[c]), cls_m2[c], cls_mu3[c]};
            for (int w = 0; w < 3; ++w)
                for (int l = 0; l <= L; l += b->lstep) lamyr.push_back(l * W + ms[w] + L);
        }
        lamys.push_back((int)lamyr.size());
        b->lam_rows = 3 * nr;
        const auto lof = [&](int y) { return y / K / W; };          // l of a table row y K
        const auto mof = [&](int y) { return y / K % W - L; };      // m of a table row y K
        for (const LamEnt &x : lament)
            lamrec.push_back(LamRec{x.hw, (uint32_t)x.lt | ((uint32_t)x.fl << 16),
                                    (uint32_t)(lof(x.y1) / b->lstep) | ((uint32_t)(nr + lof(x.y2) / b->lstep) << 8) |
                                        ((uint32_t)(2 * nr + lof(x.y3) / b->lstep) << 16)});
        b->lam_lds = glam_lds_bytes(b->lam_rows, K, max_ent);
        // backward: l-triple t stages (l1, m), (l2, m), (l3, mu) for every m;
        // its terms are the forward ones of every class, sorted by (r3, r2, r1)
        std::vector<std::vector<std::pair<int, LamRec>>> byt(nlt);
        for (int c = 0; c < b->n_m; ++c)
            for (int e = lamcs[c]; e < lamcs[c + 1]; ++e) {
                const LamEnt &x = lament[e];
                const int l1 = lof(x.y1), l2 = lof(x.y2), l3 = lof(x.y3);
                const uint32_t r1 = (uint32_t)(mof(x.y1) + l1), r2 = (uint32_t)(2 * l1 + 1 + mof(x.y2) + l2),
                               r3 = (uint32_t)(2 * l1 + 1 + 2 * l2 + 1 + mof(x.y3) + l3);
                byt[x.lt].push_back({0, LamRec{x.hw, (uint32_t)c, r1 | (r2 << 8) | (r3 << 16)}});
            }
        int maxb_ent = 0;
        for (int t = 0; t < nlt; ++t) {
            const int l3 = ltinfo[t].z, l1 = ltinfo[t].w & 255, l2 = ltinfo[t].w >> 8;
            lambys.push_back((int)lambyr.size());
            for (int m = -l1; m <= l1; ++m) lambyr.push_back(l1 * W + m + L);
            for (int m = -l2; m <= l2; ++m) lambyr.push_back(l2 * W + m + L);
            for (int m = -l3; m <= l3; ++m) lambyr.push_back(l3 * W + m + L);
            b->lamb_rows = std::max(b->lamb_rows, 2 * (l1 + l2 + l3) + 3);
            auto &v = byt[t];
            std::sort(v.begin(), v.end(), [](const std::pair<int, LamRec> &x, const std::pair<int, LamRec> &y) {
                const uint32_t a3 = x.second.b >> 16, b3 = y.second.b >> 16;
                const uint32_t a2 = (x.second.b >> 8) & 255u, b2 = (y.second.b >> 8) & 255u;
                if (a3 != b3) return a3 < b3;
                if (a2 != b2) return a2 < b2;
                return (x.second.b & 255u) != (y.second.b & 255u) ? (x.second.b & 255u) < (y.second.b & 255u)
                                                                  : x.second.a < y.second.a;
            });
            lambcs.push_back((int)lambrec.size());
            for (size_t i = 0; i < v.size(); ++i) {
                LamRec rrec = v[i].second;
                const bool n3 = i == 0 || (rrec.b >> 16) != (v[i - 1].second.b >> 16);
                const bool n2 = n3 || ((rrec.b >> 8) & 255u) != ((v[i - 1].second.b >> 8) & 255u);
                rrec.a |= (uint32_t)((n3 ? 1 : 0) | (n2 ? 2 : 0)) << 16;
                lambrec.push_back(rrec);
            }
            maxb_ent = std::max(maxb_ent, (int)v.size());
        }
        lambcs.push_back((int)lambrec.size());
        lambys.push_back((int)lambyr.size());
        b->lamb_lds = glam_lds_bytes(b->lamb_rows, K, maxb_ent);
        if (b->lam_lds > 65536) {
            (void)hipFuncSetAttribute((const void *)k_glam_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b->lam_lds);
        }
        if (std::max(b->lam_lds, b->lamb_lds) > 65536)
            (void)hipFuncSetAttribute((const void *)k_glam, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)std::max(b->lam_lds, b->lamb_lds));
        for (int k1 = 1; k1 <= K; ++k1)
            for (int k2 = 1; k2 <= k1; ++k2)
                for (int k3 = 1; k3 <= k2; ++k3) ktk.push_back(k1 | (k2 << 10) | (k3 << 20));
    }
    // the same groups padded to multiples of 8 entries (zero weight, the
    // group's first index) for the FP16 triple products
    std::vector<float> ew8;
    std::vector<uint32_t> eidx8;
    std::vector<int> gstart8;
    for (size_t g = 0; g + 1 < gstart.size(); ++g) {
        gstart8.push_back((int)ew8.size());
        const int a = gstart[g], n = gstart[g + 1] - gstart[g], n8 = (n + 7) / 8 * 8;
        for (int i = 0; i < n8; ++i) {
            ew8.push_back(i < n ? (float)ew[a + i] : 0.f);
            eidx8.push_back(eidx[i < n ? a + i : a]);
        }
    }
    gstart8.push_back((int)ew8.size());
    b->basislen = basislen;
    b->n_lt = nlt;
    b->n_groups = (int)gstart.size() - 1;

    // chunks of consecutive l-triples for k_triple, balanced by the work of
    // both stages (entries per pair + (2 l3 + 1) per output), at most
    // kChunkGroups (lt,mu3) groups each
    constexpr int kChunkGroups = 64;
    std::vector<int4> chunks;
    {
        double total = 0.0;
        std::vector<double> cost(nlt);
        for (int t = 0; t < nlt; ++t) {
            cost[t] = lt_entries[t] + 8.0 * (2 * ltinfo[t].z + 1);
            total += cost[t];
        }
        const double per = total / 512.0;
        auto gfirst = [&](int t) { return t < nlt ? lt_goff[t] : b->n_groups; };
        int lt0 = 0;
        double acc = 0.0;
        for (int t = 0; t < nlt; ++t) {
            acc += cost[t];
            // close after lt t when the chunk carries its share of the work,
            // or when lt t+1 would push it past kChunkGroups groups
            if (acc >= per || t == nlt - 1 || gfirst(t + 2) - gfirst(lt0) > kChunkGroups) {
                chunks.push_back(make_int4(lt0, t + 1, gfirst(lt0), gfirst(t + 1)));
                lt0 = t + 1;
                acc = 0.0;
            }
        }
    }
    b->nchunks = (int)chunks.size();
    b->Gmax = 0;
    b->Emax = 0;
    for (auto &c : chunks) {
        b->Gmax = std::max(b->Gmax, c.w - c.z);
        b->Emax = std::max(b->Emax, gstart[c.w] - gstart[c.z]);
    }
    std::vector<int> pairs;
    for (int k1 = 1; k1 <= K; ++k1)
        for (int k2 = 1; k2 <= k1; ++k2) pairs.push_back(k1 | (k2 << 16));
    b->npairs = (int)pairs.size();
    {
        // k_triple keeps all K shells, the chunk's weights and a batch of PB
        // pairs' T values in LDS
        int maxlds = 0;
        if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess)
            maxlds = 65536;
        const size_t fixed = sizeof(double2) * (size_t)K * nc + (sizeof(double) + sizeof(uint32_t)) * (size_t)b->Emax +
                             sizeof(int) * ((size_t)b->Gmax + 1) + 64;
        const size_t per_pair = sizeof(double2) * (size_t)b->Gmax;
        if (fixed + per_pair > (size_t)maxlds) {
            delete b;
            return tp_fail(TP_E_UNSUPPORTED, "K shells of coefficients do not fit the workgroup LDS (K too large for L)");
        }
        b->PB = (int)std::min<size_t>((size_t)b->npairs, ((size_t)maxlds - fixed) / per_pair);
        b->PB = std::min(b->PB, 256);
        b->triple_lds = fixed + per_pair * b->PB;
        if (b->triple_lds > 65536)
            (void)hipFuncSetAttribute((const void *)k_triple, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b->triple_lds);
    }

    // MFMA triple products: one workgroup per (l-triple, half of the k1
    // range), most expensive first
    std::vector<int> work;
    if (K <= kMfmaMaxK) {
        b->Lw = 2 * L + 1;
        b->EmaxLt = 0;
        for (int t = 0; t < nlt; ++t) b->EmaxLt = std::max(b->EmaxLt, lt_entries[t]);
        const int nh = K > 16 ? 2 : 1;
        int maxrb = 0;
        std::vector<std::pair<double, int>> wl;
        for (int h = 0; h < nh; ++h) {
            const int k1lo = 16 * h + 1, k1hi = std::min(16 * h + 16, K);
            const int P = k1hi * (k1hi + 1) / 2 - (k1lo - 1) * k1lo / 2;
            const int nrb = (P + 15) / 16;
            maxrb = std::max(maxrb, nrb);
            for (int t = 0; t < nlt; ++t) {
                const int n3 = 2 * ltinfo[t].z + 1;
                const double cost = 2.0 * lt_entries[t] + 8.0 * ((n3 + 7) / 8) * nrb;
                wl.push_back({-cost, t | (h << 16)});
            }
        }
        std::sort(wl.begin(), wl.end());
        for (auto &w : wl) work.push_back(w.second);
        b->nwork = (int)work.size();
        b->Prows = 16 * maxrb;
        b->mfma_lds = sizeof(double2) * 3 * (size_t)K * b->Lw + sizeof(double) * 32 * (size_t)b->Prows +
                      (sizeof(double) + sizeof(uint32_t)) * (size_t)b->EmaxLt + sizeof(int) * ((size_t)b->Lw + 1) +
                      sizeof(int) * (size_t)b->Prows + 64;
        int maxlds = 0;
        if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess)
            maxlds = 65536;
        b->EmaxLt8 = 0;
        for (int t = 0; t < nlt; ++t) {
            const int g0 = ltinfo[t].y, n3 = 2 * ltinfo[t].z + 1;
            b->EmaxLt8 = std::max(b->EmaxLt8, gstart8[g0 + n3] - gstart8[g0]);
        }
        b->t16_lds = triple16_lds(K, b->Lw, b->EmaxLt8, b->Prows);
        b->use_mfma = b->mfma_lds <= (size_t)maxlds && b->t16_lds <= (size_t)maxlds && maxrb <= 16;
        if (b->use_mfma && b->mfma_lds > 65536)
            (void)hipFuncSetAttribute((const void *)k_triple_mfma, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b->mfma_lds);
        if (b->use_mfma && b->t16_lds > 65536) {
            (void)hipFuncSetAttribute((const void *)k_triple16, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b->t16_lds);
            (void)hipFuncSetAttribute((const void *)k_triple16_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)b->t16_lds);
        }
    }

    // sorted l-triples for k_triple_sym: each ordered l-triple t maps to the
    // sorted one ts (l1 >= l2 >= l3) by a permutation sigma, l_i = L_sigma(i);
    // only for the library's own (permutation-symmetric) 3j weights
    std::vector<int> perms, gorder, symoff;
    std::vector<SymItem> sym_work;
    std::vector<uint32_t> symtab;
    if (!hw && K <= 32 && nlt > 0 && nlt < 2048) {   // sym_tsx: 11 bits of l-triple index
        std::unordered_map<int, int> index_of;   // l1 | l2 << 8 | l3 << 16 -> lt
        for (int t = 0; t < nlt; ++t)
            index_of[(ltinfo[t].w & 0xffff) | (ltinfo[t].z << 16)] = t;
        perms.assign((size_t)nlt * 8, 0);
        gorder.assign((size_t)b->n_groups, 0);
        bool ok = true;
        std::vector<int> sorted_ts;
        for (int t = 0; t < nlt && ok; ++t) {
            const int l[3] = {ltinfo[t].w & 0xff, (ltinfo[t].w >> 8) & 0xff, ltinfo[t].z};
            int Ls[3] = {l[0], l[1], l[2]};
            std::sort(Ls, Ls + 3, [](int a, int c) { return a > c; });
            const auto it = index_of.find(Ls[0] | (Ls[1] << 8) | (Ls[2] << 16));
            if (it == index_of.end()) { ok = false; break; }
            const int ts = it->second;
            int sig[3];
            bool used[3] = {false, false, false};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    if (!used[j] && Ls[j] == l[i]) { sig[i] = j; used[j] = true; break; }
            int pidx = -1;
            for (int p = 0; p < 6; ++p)
                if (kSymPerm[p][0] == sig[0] && kSymPerm[p][1] == sig[1] && kSymPerm[p][2] == sig[2]) pidx = p;
            int &cnt = perms[(size_t)ts * 8];
            if (cnt == 0) sorted_ts.push_back(ts);
            if (cnt >= 6 || pidx < 0) { ok = false; break; }
            perms[(size_t)ts * 8 + 1 + cnt] = t | (pidx << 24);
            ++cnt;
        }
        int emax_lt = 0;
        for (int t = 0; t < nlt; ++t) emax_lt = std::max(emax_lt, lt_entries[t]);
        const size_t lds = triple_sym_lds(K, 2 * L + 1, emax_lt);
        const size_t lds16 = triple_sym_lds(K, 2 * L + 1, emax_lt, kSymWaves, true);
        const size_t lds_b = triple_sym_lds(K, 2 * L + 1, emax_lt, kSymWavesBatch, true);
        int maxlds = 0;
        if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess)
            maxlds = 65536;
        if (ok && lds <= (size_t)maxlds) {
            const int nh = K > 16 ? 2 : 1;
            std::vector<std::pair<double, int>> wl;
            for (int ts : sorted_ts) {
                const int g0 = ltinfo[ts].y, n3 = 2 * ltinfo[ts].z + 1;
                // the groups of a chunk run together: largest first, so that a chunk's are alike
                std::vector<std::pair<int, int>> ge;
                double s1 = 0.0;
                for (int u = 0; u < n3; ++u) {
                    const int E = gstart[g0 + u + 1] - gstart[g0 + u];
                    ge.push_back({-E, u});
                }
                std::stable_sort(ge.begin(), ge.end());
                for (int u = 0; u < n3; ++u) gorder[(size_t)g0 + u] = ge[u].second;
                for (int c0 = 0; c0 < n3; c0 += kSymWaves) s1 += (-ge[c0].first + 3) / 4;   // slowest wave
                const double s2 = ((n3 + 1) / 2) * 2.0 * (nh > 1 ? 2 : 1);
                const double cost = 2.0 * s1 + s2 + 0.25 * perms[(size_t)ts * 8];
                b->h_sym_cost.push_back({ts, 2.0 * s1 + 0.25 * perms[(size_t)ts * 8]});
                // with l1 = l2 the items h1 < h2 are stored by their mirror images
                const bool sym12 = (ltinfo[ts].w & 0xff) == ((ltinfo[ts].w >> 8) & 0xff);
                for (int h1 = 0; h1 < nh; ++h1)
                    for (int h2 = 0; h2 < nh; ++h2)
                        if (!(sym12 && h1 < h2)) wl.push_back({-cost, ts | (h1 << 16) | (h2 << 20)});
            }
            std::stable_sort(wl.begin(), wl.end());
            // the work items: each l-triple's record with the item's halves
            b->h_sym_item.assign((size_t)nlt, SymItem{});
            for (int t = 0; t < nlt; ++t) {
                const int g0 = ltinfo[t].y, n3 = 2 * ltinfo[t].z + 1;
                SymItem &it = b->h_sym_item[(size_t)t];
                it.tsx = sym_tsx(t, 0, 0, ltinfo[t].w & 0xff, (ltinfo[t].w >> 8) & 0xff, ltinfo[t].z);
                it.e0 = gstart[g0];
                it.ne = gstart[g0 + n3] - gstart[g0];
                it.goff = g0;
                it.K = K;
            }
            for (auto &w : wl) {
                SymItem it = b->h_sym_item[(size_t)(w.second & 0xffff)];
                it.tsx |= ((w.second >> 16) & 15) << 11 | ((w.second >> 20) & 15) << 12;
                sym_work.push_back(it);
            }
            // per (permutation, K1 half, K2 half, K3 block): the tile entries
            // V[K1][K2][K3] that are As[lt][kt] for the sorted k-triple
            // kt = (k1, k2, k3), k_i = K_sigma(i), in kt order
            for (int p = 0; p < 6; ++p)
                for (int h1 = 0; h1 < 2; ++h1)
                    for (int h2 = 0; h2 < 2; ++h2)
                        for (int cb = 0; cb < 2; ++cb) {
                            symoff.push_back((int)symtab.size());
                            std::vector<std::pair<int, uint32_t>> ents;
                            for (int a = 0; a < 16; ++a)
                                for (int c = 0; c < 16; ++c)
                                    for (int d = 0; d < 16; ++d) {
                                        const int Kv[3] = {16 * h1 + a + 1, 16 * h2 + c + 1, 16 * cb + d + 1};
                                        if (Kv[0] > K || Kv[1] > K || Kv[2] > K) continue;
                                        const int k1 = Kv[kSymPerm[p][0]], k2 = Kv[kSymPerm[p][1]],
                                                  k3 = Kv[kSymPerm[p][2]];
                                        if (k1 < k2 || k2 < k3) continue;
                                        const int kt = (k1 - 1) * k1 * (k1 + 1) / 6 + (k2 - 1) * k2 / 2 + (k3 - 1);
                                        ents.push_back({kt, (uint32_t)(a * 256 + c * 16 + d) | ((uint32_t)kt << 12)});
                                    }
                            std::sort(ents.begin(), ents.end());
                            for (auto &x : ents) symtab.push_back(x.second);
                        }
            symoff.push_back((int)symtab.size());
            b->nsym_work = (int)sym_work.size();
            b->sym_lds = lds;
            b->sym16_lds = lds16;
            b->sym_lds_b = lds_b;
            b->use_sym = true;
            if (lds > 65536)
                (void)hipFuncSetAttribute((const void *)k_triple_sym, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds);
            if (lds16 > 65536)
                (void)hipFuncSetAttribute((const void *)k_triple_sym16, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)lds16);
            if (lds_b > 65536)
                (void)hipFuncSetAttribute((const void *)k_triple_sym_batch,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_b);
            if (b->Lw == 0) b->Lw = 2 * L + 1;
        }
    }

    // padded dims; H and Bm
    b->R_pad = round_up(b->R, G2_BM);
    b->nlt_pad = round_up(nlt, 16);
    b->kl_pad = round_up(b->klength, KL_ALIGN);
    {   // k_glam: the columns in equal blocks of at most kLamMaxThreads
        const int64_t nblk = (b->kl_pad + kLamMaxThreads - 1) / kLamMaxThreads;
        b->lam_nt = (int)round_up((b->kl_pad + nblk - 1) / nblk, 64);
    }
    b->nm_pad64 = round_up(b->n_m, G1_BM);
    b->nm_pad16 = round_up(b->n_m, 16);
    b->n_rt = (int)(b->R_pad / G2_BM);
    b->n_ct = (int)(b->kl_pad / G2_BN);
    b->ldk = (int)round_up(nlt, 64);
    b->R_pad16 = round_up(b->R, H_BM);
    b->n_rt16 = (int)(b->R_pad16 / H_BM);
    b->n_ct16 = (int)(b->kl_pad / H_BN);
    std::vector<double> H((size_t)b->nlt_pad * b->nm_pad64, 0.0);
    for (int t = 0; t < nlt; ++t)
        for (int c = 0; c < b->n_m; ++c) H[(size_t)t * b->nm_pad64 + c] = Hcols[t][c];
    std::vector<double> Bm((size_t)b->nm_pad16 * b->R_pad, 0.0);
    for (int c = 0; c < b->n_m; ++c)
        for (int jb = 0; jb < 2 * N; ++jb)
            for (int ia = 0; ia < N; ++ia) {
                // B[r,j] = cos(m2 a + mu3 b), alpharange / alpharange_2pi (correlations.jl:141,247-256)
                const double a = (ia + 0.5) * tph::kPi / N, be = (jb + 0.5) * tph::kPi / N;
                Bm[(size_t)c * b->R_pad + ia + (size_t)jb * N] = std::cos(cls_m2[c] * a + cls_mu3[c] * be);
            }

    // separated-grid tables (k_c3sep)
    b->n_m2 = L / mstep + 1;
    b->jpad = (int)round_up(2 * N, 16);
    b->sep_kb = (int)round_up(2 * b->n_m2, 4);
    b->njc = b->jpad / 16;
    std::vector<int> m2start(b->n_m2 + 1, b->n_m);
    for (int c = b->n_m - 1; c >= 0; --c) m2start[cls_m2[c] / mstep] = c;
    for (int i = b->n_m2 - 1; i >= 0; --i)
        if (m2start[i] > m2start[i + 1]) m2start[i] = m2start[i + 1];
    int max_cls = 0;
    for (int i = 0; i < b->n_m2; ++i) max_cls = std::max(max_cls, m2start[i + 1] - m2start[i]);
    b->use_sep = N <= 32 && b->sep_kb <= kSepMaxKB && max_cls <= 24;
    b->sep_wide = !b->use_sep && N <= 32 && b->sep_kb <= kSepWideMaxKB;
    if (b->sep_wide && c3sep_lds_bytes(b->sep_kb) > 65536) {
        const int lb = (int)c3sep_lds_bytes(b->sep_kb);
        (void)hipFuncSetAttribute((const void *)k_c3sep<1, true, false, false, kSepWideQB>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lb);
        (void)hipFuncSetAttribute((const void *)k_c3sep<0, true, false, false, kSepWideQB>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lb);
        (void)hipFuncSetAttribute((const void *)k_c3sep<1, false, false, false, kSepWideQB>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lb);
        (void)hipFuncSetAttribute((const void *)k_c3sep<0, true, true, false, kSepWideQB>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lb);
    }
    b->fold = b->use_sep && !b->lam && N % 2 == 0;
    b->njf = (int)round_up(N, 16) / 16;
    std::vector<double> TA((size_t)2 * b->n_m * b->jpad, 0.0);
    for (int c = 0; c < b->n_m; ++c)
        for (int j = 0; j < 2 * N; ++j) {
            const double be = (j + 0.5) * tph::kPi / N;
            TA[(size_t)c * b->jpad + j] = std::cos(cls_mu3[c] * be);
            TA[(size_t)(b->n_m + c) * b->jpad + j] = std::sin(cls_mu3[c] * be);
        }
    const int nip = (int)round_up(N, 16);
    std::vector<double> TB((size_t)nip * b->sep_kb, 0.0);
    for (int i = 0; i < N; ++i) {
        const double a = (i + 0.5) * tph::kPi / N;
        for (int m2i = 0; m2i < b->n_m2; ++m2i) {
            TB[(size_t)i * b->sep_kb + 2 * m2i] = std::cos(mstep * m2i * a);
            TB[(size_t)i * b->sep_kb + 2 * m2i + 1] = -std::sin(mstep * m2i * a);
        }
    }

    std::vector<double> logtab(kLogTabSize);
    tp_fastlog_table(logtab.data());

    int rc = 0;
    auto up = [&](auto **dptr, const auto &vec) -> int {
        using T = typename std::remove_reference<decltype(vec)>::type::value_type;
        if (int r2 = dev_alloc(dptr, vec.size())) return r2;
        hipError_t e = hipMemcpy((void *)*dptr, vec.data(), sizeof(T) * vec.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return tp_fail(TP_E_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return 0;
    };
    if ((rc = up(&b->d_H, H)) || (rc = up(&b->d_Bm, Bm)) || (rc = up(&b->d_ew, ew)) || (rc = up(&b->d_eidx, eidx)) ||
        (rc = up(&b->d_gstart, gstart)) || (rc = up(&b->d_ltinfo, ltinfo)) || (rc = up(&b->d_chunks, chunks)) ||
        (rc = up(&b->d_pairs, pairs)) || (rc = up(&b->d_sean0, b->sean0)) || (rc = up(&b->d_coff, b->coff)) ||
        (b->use_mfma && (rc = up(&b->d_work, work))) ||
        (b->use_sym && ((rc = up(&b->d_sym_work, sym_work)) || (rc = up(&b->d_perms, perms)) ||
                        (rc = up(&b->d_gorder, gorder)) || (rc = up(&b->d_symtab, symtab)) ||
                        (rc = up(&b->d_symoff, symoff)))) ||
        (rc = up(&b->d_ew8, ew8)) || (rc = up(&b->d_eidx8, eidx8)) ||
        (rc = up(&b->d_gstart8, gstart8)) || (rc = up(&b->d_m2start, m2start)) || (rc = up(&b->d_TA, TA)) ||
        (rc = up(&b->d_TB, TB)) || (rc = up(&b->d_logtab, logtab))) {
        tp_basis_destroy(b);
        return rc;
    }
    b->h_ltinfo = ltinfo;
    if ((rc = dev_alloc(&b->d_As, (size_t)b->nlt_pad * b->kl_pad)) ||
        (rc = dev_alloc(&b->d_G, (size_t)b->nm_pad64 * b->kl_pad)) ||
        (rc = dev_alloc(&b->d_part, (size_t)std::max(b->n_rt, b->njc) * 2 * b->kl_pad)) ||
        (rc = dev_alloc(&b->d_out, 8)) ||
        (rc = dev_alloc(&b->d_raw, (size_t)K * LMAX * LMAX)) || (rc = dev_alloc(&b->d_coef, (size_t)K * b->n_c)) ||
        (rc = dev_alloc(&b->d_shsc, (size_t)2 * K)) || (rc = dev_alloc(&b->d_nrm2, (size_t)K * kNrmLd))) {
        tp_basis_destroy(b);
        return rc;
    }
    if (b->lam && ((rc = up(&b->d_lamY, lamY)) || (rc = up(&b->d_ktk, ktk)) || (rc = up(&b->d_lament, lamrec)) ||
                   (rc = up(&b->d_lamcs, lamcs)) || (rc = up(&b->d_lamyr, lamyr)) || (rc = up(&b->d_lamys, lamys)) ||
                   (rc = up(&b->d_lambent, lambrec)) || (rc = up(&b->d_lambcs, lambcs)) ||
                   (rc = up(&b->d_lambyr, lambyr)) || (rc = up(&b->d_lambys, lambys)))) {
        tp_basis_destroy(b);
        return rc;
    }
    // padding of As (l-triple rows >= n_lt, columns >= klength) must read as
    // zero; so must the rows of G past the classes, which k_glam leaves alone
    hipError_t e = hipMemset(b->d_As, 0, sizeof(double) * (size_t)b->nlt_pad * b->kl_pad);
    if (e == hipSuccess) e = hipMemset(b->d_G, 0, sizeof(double) * (size_t)b->nm_pad64 * b->kl_pad);
    if (e != hipSuccess) {
        tp_basis_destroy(b);
        return tp_fail(TP_E_HIP, "hipMemset failed");
    }
    b->h_H = std::move(H);
    b->h_Bm = std::move(Bm);
    b->h_lt_cnt = lt_entries;
    for (auto &t : ltinfo) b->h_lt_l3.push_back(t.z);
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


int tp_basis_set_precision(tp_basis *b, int mode)
{
    if (!b) return tp_fail(TP_E_ARG, "null basis");
    if (mode != TP_PREC_FP64 && mode != TP_PREC_FP16X3 && mode != TP_PREC_BF16 && mode != TP_PREC_TRIPLE16 &&
        mode != TP_PREC_FP16X3_T64)
        return tp_fail(TP_E_ARG, "unknown precision mode");
    if (mode == TP_PREC_FP16X3_T64 && !b->use_sym)
        return tp_fail(TP_E_UNSUPPORTED, "TP_PREC_FP16X3_T64 needs the sorted l-triple tables (the library's own "
                                         "weights, K <= 32)");
    if (int rc = check_ctx(b->ctx)) return rc;
    if (mode != TP_PREC_FP64 && (b->K > 64 || b->n_c > 1024))
        return tp_fail(TP_E_UNSUPPORTED, "mixed-precision contraction supports K <= 64 and L <= 44");
    if (mode != TP_PREC_FP64 && mode != TP_PREC_FP16X3_T64 && b->lam)
        return tp_fail(TP_E_UNSUPPORTED, "on the curved Ewald sphere (lambda dq != 0) the precision modes are FP64 "
                                         "and FP16X3_T64");
    if (mode == TP_PREC_FP16X3_T64 && !b->d_Hhi && !b->lam) {   // lambda dq != 0: G from k_glam in FP64
        // H[class][lt] (FP64 on the host) scaled by a power of two into
        // [2^13, 2^14) and split into FP16 hi/lo for k_hgemm_g
        if (b->n_m > 16 * kHgMaxRb || b->L / 2 >= kNrmLd)
            return tp_fail(TP_E_UNSUPPORTED, "TP_PREC_FP16X3_T64 supports at most 192 classes (L <= 20)");
        // the column scales come from the l-block norms that the work items of
        // the l-triples (l, l, l) write: every even l <= L needs one
        for (int l = 0; l <= b->L; l += 2) {
            bool found = false;
            for (const int4 &t : b->h_ltinfo)
                found = found || (t.z == l && (t.w & 0xff) == l && ((t.w >> 8) & 0xff) == l);
            if (!found) return tp_fail(TP_E_UNSUPPORTED, "TP_PREC_FP16X3_T64 needs the l-triple (l, l, l) of every l");
        }
        const int nrb = (b->n_m + 15) / 16;
        double hmax = 0.0;
        for (int64_t t = 0; t < b->n_lt; ++t)
            for (int c = 0; c < b->n_m; ++c) hmax = std::max(hmax, std::fabs(b->h_H[(size_t)t * b->nm_pad64 + c]));
        const double sH = hmax > 0.0 ? std::ldexp(1.0, 13 - std::ilogb(hmax)) : 1.0;
        const int ldh = (int)round_up(b->n_lt, 128);
        std::vector<uint16_t> hh((size_t)kHgRows * ldh, 0), hl(hh.size(), 0);
        for (int c = 0; c < b->n_m; ++c)
            for (int64_t t = 0; t < b->n_lt; ++t) {
                const double x = b->h_H[(size_t)t * b->nm_pad64 + c] * sH;
                const size_t i = (size_t)c * ldh + t;
                hh[i] = host_f16(x);
                hl[i] = host_f16(x - host_f16_val(hh[i]));
            }
        uint16_t *dh = nullptr, *dl = nullptr;
        if (int rc = dev_alloc(&dh, hh.size())) return rc;
        if (int rc = dev_alloc(&dl, hl.size())) {
            (void)hipFree(dh);
            return rc;
        }
        if (hipMemcpy(dh, hh.data(), sizeof(uint16_t) * hh.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dl, hl.data(), sizeof(uint16_t) * hl.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(dh);
            (void)hipFree(dl);
            return tp_fail(TP_E_HIP, "hipMemcpy failed");
        }
        b->d_Hhi = dh;
        b->d_Hlo = dl;
        b->H_inv = 1.0 / sH;
        b->nrb16 = nrb;
        b->ldh16 = ldh;
    }
    if ((mode == TP_PREC_FP16X3 || mode == TP_PREC_BF16) && !b->d_Fhi) {
        // F[r][lt] = sum_class Bm[class][r] H[lt][class] in FP64, then scaled by a
        // power of two into [2^13, 2^14) and split into FP16 hi/lo and BF16
        const int64_t nlt = b->n_lt;
        std::vector<double> F((size_t)b->R_pad16 * b->ldk, 0.0);
        double fmax = 0.0;
        for (int64_t r = 0; r < b->R; ++r)
            for (int64_t t = 0; t < nlt; ++t) {
                double s = 0.0;
                for (int c = 0; c < b->n_m; ++c)
                    s += b->h_Bm[(size_t)c * b->R_pad + r] * b->h_H[(size_t)t * b->nm_pad64 + c];
                F[(size_t)r * b->ldk + t] = s;
                fmax = std::max(fmax, std::fabs(s));
            }
        const double sF = fmax > 0.0 ? std::ldexp(1.0, 13 - std::ilogb(fmax)) : 1.0;
        b->F_inv = 1.0 / sF;
        std::vector<uint16_t> fh(F.size()), fl(F.size()), fb(F.size());
        for (size_t i = 0; i < F.size(); ++i) {
            const double x = F[i] * sF;
            fh[i] = host_f16(x);
            fl[i] = host_f16(x - host_f16_val(fh[i]));
            fb[i] = host_bf16(x);
        }
        int rc = 0;
        auto up16 = [&](uint16_t **d, const std::vector<uint16_t> &v) -> int {
            if (int r2 = dev_alloc(d, v.size())) return r2;
            if (hipMemcpy(*d, v.data(), sizeof(uint16_t) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
                return tp_fail(TP_E_HIP, "hipMemcpy failed");
            return 0;
        };
        if ((rc = up16(&b->d_Fhi, fh)) || (rc = up16(&b->d_Flo, fl)) || (rc = up16(&b->d_Fbf, fb)) ||
            (rc = dev_alloc(&b->d_Ahi, (size_t)b->kl_pad * b->ldk)) ||
            (rc = dev_alloc(&b->d_Alo, (size_t)b->kl_pad * b->ldk)) || (rc = dev_alloc(&b->d_Ainv, b->kl_pad)))
            return rc;
    }
    b->prec = mode;
    return 0;
}

int tp_basis_get_precision(const tp_basis *b)
{
    return b ? b->prec : TP_E_ARG;
}

int tp_c3(tp_ctx *ctx, const tp_basis *bc, const double *coeff, int64_t ld, double *c3_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !c3_out) return tp_fail(TP_E_ARG, "null argument");
    tp_basis *b = const_cast<tp_basis *>(bc);
    if (b->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
    if (!b->d_c3)
        if (int rc = dev_alloc(&b->d_c3, (size_t)b->klength * b->R)) return rc;
    if (int rc = upload_coeff(ctx, b, coeff, ld, b->K, b->d_coef, b->d_raw)) return rc;
    if (int rc = enqueue_eval(ctx, b, nullptr, b->d_coef, nullptr, nullptr, 0, b->d_c3, nullptr, mc_none())) return rc;
    TP_HIP(hipMemcpyAsync(c3_out, b->d_c3, sizeof(double) * (size_t)b->klength * b->R, hipMemcpyDeviceToHost,
                          ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

void tp_ref_destroy(tp_ref *r)
{
    if (!r) return;
    if (r->b && r->b->ctx) (void)hipSetDevice(r->b->ctx->device);
    if (r->d_qhat) (void)hipFree(r->d_qhat);
    if (r->d_qfold) (void)hipFree(r->d_qfold);
    if (r->d_qsum) (void)hipFree(r->d_qsum);
    if (r->d_tf) (void)hipFree(r->d_tf);
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
    const int64_t R = b->R;
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
        if (!(s > 0.0) || !std::isfinite(s)) {
            char msg[192];
            std::snprintf(msg, sizeof(msg),
                          "histogram slice (k1,k2,k3) = (%d,%d,%d) sums to %g: no normalised form; clamp it as "
                          "loadHistograms does (tp_load_histograms)", k1, k2, k3, s);
            return tp_fail(TP_E_ARG, msg);
        }
        double t = 0.0;
        for (int64_t r = 0; r < R; ++r) {
            qhat[(size_t)i * R + r] = q[r] / s;
            t += q[r] / s;
        }
        qsum[i] = t;
        // tripletFactor (datagen.jl:166-174)
        tf[i] = (k1 == k2 && k2 == k3) ? 1.0 : ((k1 == k2 || k2 == k3 || k1 == k3) ? 3.0 : 6.0);
    }
    tp_ref *r = new tp_ref();
    r->b = b;
    r->klo = k3lo;
    r->khi = k3hi;
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
    if (b->fold) {
        // each orbit {(i, j), (i, j + N), (N-1-i, 2N-1-j), (N-1-i, N-1-j)} of
        // i < N/2, j < N summed into bin i + (N/2) j of the folded column
        const int N = b->N, Nh = N / 2;
        const int64_t RF = (int64_t)Nh * N;
        std::vector<double> qf((size_t)(ncols * RF));
        for (int64_t c = 0; c < ncols; ++c) {
            const double *q = qhat.data() + (size_t)(c * R);
            for (int j = 0; j < N; ++j)
                for (int i = 0; i < Nh; ++i)
                    qf[(size_t)(c * RF + i + (int64_t)Nh * j)] =
                        q[i + (int64_t)N * j] + q[i + (int64_t)N * (j + N)] + q[(N - 1 - i) + (int64_t)N * (2 * N - 1 - j)] +
                        q[(N - 1 - i) + (int64_t)N * (N - 1 - j)];
        }
        if (dev_alloc(&r->d_qfold, qf.size())) {
            tp_ref_destroy(r);
            return TP_E_OOM;
        }
        if (hipMemcpy(r->d_qfold, qf.data(), sizeof(double) * qf.size(), hipMemcpyHostToDevice) != hipSuccess) {
            tp_ref_destroy(r);
            return tp_fail(TP_E_HIP, "histogram upload failed");
        }
    }
    *out = r;
    return 0;
}

int tp_energy(tp_ctx *ctx, const tp_basis *bc, const tp_ref *r, const double *coeff, int64_t ld, int measure,
              double *E_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !r || !E_out) return tp_fail(TP_E_ARG, "null argument");
    if (measure != 0) {   // correlations.jl:210
        *E_out = 0.0;
        return 0;
    }
    tp_basis *b = const_cast<tp_basis *>(bc);
    if (b->ctx != ctx || r->b != b) return tp_fail(TP_E_STATE, "objects belong to another context/basis");
    if (int rc = upload_coeff(ctx, b, coeff, ld, b->K, b->d_coef, b->d_raw)) return rc;
    McArgs mc = mc_none();
    mc.seqp = reinterpret_cast<uint64_t *>(ctx->d_res + 1);
    mc.seq = ++ctx->res_seq;
    if (int rc = enqueue_eval(ctx, b, r, b->d_coef, nullptr, nullptr, 0, nullptr, ctx->d_res, mc)) return rc;
    if (int rc = wait_result(ctx, mc.seq)) return rc;
    *E_out = ctx->h_res[0];
    record_timing(ctx);
    return 0;
}

// The energy with every per-evaluation buffer carved from the static scratch
// array: the coefficient staging and compact copy, As, G, the energy partial
// sums, the shell scales and norms and, in mixed precision, the 16-bit As.
// No allocation, no memset.
int tp_energy_noalloc(tp_ctx *ctx, const tp_basis *bc, const tp_ref *r, const double *coeff, int64_t ld,
                      int measure, double *E_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !r || !E_out || !coeff) return tp_fail(TP_E_ARG, "null argument");
    if (measure != 0) {   // correlations.jl:210
        *E_out = 0.0;
        return 0;
    }
    const tp_basis *b = bc;
    if (b->ctx != ctx || r->b != b) return tp_fail(TP_E_STATE, "objects belong to another context/basis");
    if (ld < (int64_t)b->LMAX * b->LMAX) return tp_fail(TP_E_ARG, "coefficient vector shorter than LMAX^2");
    if (!ctx->scratch) return tp_fail(TP_E_STATE, "no scratch address for this device");
    const bool mixed16 = b->prec == TP_PREC_FP16X3 || b->prec == TP_PREC_BF16;   // 16-bit copies of As
    const int64_t nrt = std::max<int64_t>(std::max(b->n_rt, b->njc), b->n_rt16);
    const size_t want[] = {
        sizeof(double2) * (size_t)b->K * ld,                  // 0 raw coefficients
        sizeof(double2) * (size_t)b->K * b->n_c,              // 1 compact coefficients
        sizeof(double) * (size_t)b->nlt_pad * b->kl_pad,      // 2 As
        sizeof(double) * (size_t)b->nm_pad64 * b->kl_pad,     // 3 G
        sizeof(double) * (size_t)nrt * 2 * b->kl_pad,         // 4 energy partial sums
        sizeof(double) * (size_t)2 * b->K,                    // 5 shell scales and norms
        mixed16 ? sizeof(uint16_t) * (size_t)b->kl_pad * b->ldk : 0,   // 6 As hi
        mixed16 ? sizeof(uint16_t) * (size_t)b->kl_pad * b->ldk : 0,   // 7 As lo
        mixed16 ? sizeof(double) * (size_t)b->kl_pad : 0,              // 8 1 / column scales
        sizeof(double) * (size_t)b->K * kNrmLd,               // 9 l-block norms (FP16X3_T64)
    };
    constexpr int nbuf = sizeof(want) / sizeof(want[0]);
    size_t off[nbuf], total = 0;
    for (int i = 0; i < nbuf; ++i) {
        off[i] = total;
        total += (want[i] + 255) / 256 * 256;
    }
    if (total > kNoAllocScratchBytes) {
        char msg[200];
        std::snprintf(msg, sizeof(msg), "this basis needs %zu bytes of scratch; tp_energy_noalloc has %zu "
                      "(use tp_energy)", total, kNoAllocScratchBytes);
        return tp_fail(TP_E_UNSUPPORTED, msg);
    }
    char *base = static_cast<char *>(ctx->scratch);
    double2 *raw = reinterpret_cast<double2 *>(base + off[0]);
    double2 *coef = reinterpret_cast<double2 *>(base + off[1]);
    EvalScratch sc{reinterpret_cast<double *>(base + off[2]),   reinterpret_cast<double *>(base + off[3]),
                   reinterpret_cast<double *>(base + off[4]),   reinterpret_cast<double *>(base + off[8]),
                   reinterpret_cast<double *>(base + off[5]),   reinterpret_cast<uint16_t *>(base + off[6]),
                   reinterpret_cast<uint16_t *>(base + off[7]), reinterpret_cast<double *>(base + off[9])};
    hipStream_t st = ctx->stream;
    // the array holds the last call's values: As and G padding must read as zero
    // (As, G and the partial sums are adjacent; so are the two 16-bit As)
    const int64_t nz = (int64_t)((off[4] + want[4] - off[2] + 7) / 8);
    hipLaunchKernelGGL(k_zero64, dim3((unsigned)std::min<int64_t>((nz + 255) / 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<double *>(base + off[2]), nz);
    if (mixed16) {
        const int64_t n16 = (int64_t)((off[7] + want[7] - off[6]) / 8);
        hipLaunchKernelGGL(k_zero64, dim3((unsigned)std::min<int64_t>((n16 + 255) / 256, 4096)), dim3(256), 0, st,
                           reinterpret_cast<double *>(base + off[6]), n16);
    }
    TP_HIP(hipMemcpyAsync(raw, coeff, want[0], hipMemcpyHostToDevice, st));
    const int64_t n = (int64_t)b->K * b->n_c;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, raw, ld, b->K, b->d_sean0, b->n_c,
                       coef);
    TP_HIP(hipGetLastError());
    if (int rc = enqueue_eval(ctx, b, r, coef, nullptr, nullptr, 0, nullptr, ctx->d_res, mc_none(), false, &sc))
        return rc;
    TP_HIP(hipStreamSynchronize(st));
    *E_out = ctx->h_res[0];
    record_timing(ctx);
    return 0;
}

// ---------------------------------------------------------------------------
// Energy gradient (no counterpart in the reference; see tp_grad.hpp)
// ---------------------------------------------------------------------------
static int grad_setup(tp_ctx *ctx, tp_basis *b)
{
    if (b->grad_ready) return 0;
    if (!(b->use_sep || (b->lam && b->sep_wide)) || !b->use_mfma)
        return tp_fail(TP_E_UNSUPPORTED, "the energy gradient needs N <= 32, L <= 22 and K <= 27");
    b->gLw = 2 * b->L + 1;
    b->gEmax = b->EmaxLt;
    b->bwd_lds = triple_bwd_lds_bytes(b->K, b->gLw, b->klength, b->gEmax);
    int maxlds = 0;
    if (hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) != hipSuccess)
        maxlds = 65536;
    if (b->bwd_lds > (size_t)maxlds)
        return tp_fail(TP_E_UNSUPPORTED, "the energy gradient's triple-product workgroup does not fit the LDS");
    if (b->bwd_lds > 65536)
        TP_HIP(hipFuncSetAttribute((const void *)k_triple_bwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)b->bwd_lds));
    b->nlt_pad64 = (int)round_up(b->n_lt, 64);
    const int nlt = (int)b->n_lt, nmp = (int)b->nm_pad64, njc = b->njc, K = b->K;
    // H^T for each of the njc row blocks of dE/dG (lambda dq = 0; on the
    // curved Ewald sphere the blocks are summed and k_glam runs backwards)
    std::vector<double> HT4(b->lam ? 0 : (size_t)njc * nmp * b->nlt_pad64, 0.0);
    if (!b->lam)
        for (int jc = 0; jc < njc; ++jc)
            for (int c = 0; c < nmp; ++c)
                for (int t = 0; t < nlt; ++t)
                    HT4[((size_t)jc * nmp + c) * b->nlt_pad64 + t] = b->h_H[(size_t)t * b->nm_pad64 + c];
    if (b->lam && c3sep_lds_bytes(b->sep_kb) > 65536)
        TP_HIP(hipFuncSetAttribute((const void *)k_c3sep_bwd<kSepWideQB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)c3sep_lds_bytes(b->sep_kb)));
    // reduction lists: for each even l, the (l-triple, slot) blocks whose slot has order l
    const int nl = b->L / 2 + 1;
    std::vector<int> lstart(nl + 1, 0), llist;
    for (int li = 0; li < nl; ++li) {
        lstart[li] = (int)llist.size();
        for (int t = 0; t < nlt; ++t) {
            const int4 q = b->h_ltinfo[t];
            const int ls[3] = {q.w & 0xff, (q.w >> 8) & 0xff, q.z};
            for (int sl = 0; sl < 3; ++sl)
                if (ls[sl] == 2 * li) llist.push_back(t * 3 + sl);
        }
    }
    lstart[nl] = (int)llist.size();
    // backward work order: per mu3 the pair sums (entries) and the k3 / slot sums
    std::vector<std::pair<int64_t, int>> cost(nlt);
    const int np = K * (K + 1) / 2;
    for (int t = 0; t < nlt; ++t) {
        const int4 q = b->h_ltinfo[t];
        cost[t] = {-((int64_t)b->h_lt_cnt[t] * (np + 2 * K) + (int64_t)(2 * q.z + 1) * 3 * np), t};
    }
    std::sort(cost.begin(), cost.end());
    std::vector<int> order(nlt);
    for (int t = 0; t < nlt; ++t) order[t] = cost[t].second;
    int rc = 0;
    if ((!b->lam && (rc = dev_alloc(&b->d_HT4, HT4.size()))) ||
        (b->lam && (rc = dev_alloc(&b->d_gG, (size_t)b->n_m * b->kl_pad))) ||
        (rc = dev_alloc(&b->d_gGp, (size_t)njc * nmp * b->kl_pad)) ||
        (rc = dev_alloc(&b->d_gA, (size_t)b->nlt_pad64 * b->kl_pad)) ||
        (rc = dev_alloc(&b->d_gpart, (size_t)nlt * 3 * K * b->gLw)) || (rc = dev_alloc(&b->d_grad, (size_t)K * b->n_c)) ||
        (rc = dev_alloc(&b->d_lstart, lstart.size())) || (rc = dev_alloc(&b->d_bwd_order, order.size())) ||
        (rc = dev_alloc(&b->d_llist, llist.size())))
        return rc;
    if (!b->lam) TP_HIP(hipMemcpy(b->d_HT4, HT4.data(), sizeof(double) * HT4.size(), hipMemcpyHostToDevice));
    // rows of dE/dAs past the l-triples (read by no kernel; k_glam writes n_lt rows) stay 0
    TP_HIP(hipMemset(b->d_gA, 0, sizeof(double) * (size_t)b->nlt_pad64 * b->kl_pad));
    TP_HIP(hipMemset(b->d_gGp, 0, sizeof(double) * (size_t)njc * nmp * b->kl_pad));   // rows / columns never written stay 0
    TP_HIP(hipMemcpy(b->d_lstart, lstart.data(), sizeof(int) * lstart.size(), hipMemcpyHostToDevice));
    TP_HIP(hipMemcpy(b->d_bwd_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
    TP_HIP(hipMemcpy(b->d_llist, llist.data(), sizeof(int) * llist.size(), hipMemcpyHostToDevice));
    b->grad_ready = true;
    return 0;
}

// dE/dc of the evaluation just enqueued (FP64, with d_G and the energy
// partials of k_c3sep still in place) into b->d_grad.
static int enqueue_grad(tp_ctx *ctx, tp_basis *b, const tp_ref *r, const double2 *coef)
{
    hipStream_t st = ctx->stream;
    const int ncols = (int)r->ncols;
    const int64_t tap = (int64_t)b->n_m * b->jpad;
    // the forward left its partial sums in njf rows (folded grid) or njc
    const int fwd_rows = b->fold && r->d_qfold ? b->njf : b->njc;
    const dim3 gbw((unsigned)((ncols + 15) / 16), (unsigned)b->njc);
    if (b->use_sep)
        hipLaunchKernelGGL(k_c3sep_bwd<6>, gbw, dim3(512), c3sep_lds_bytes(b->sep_kb), st, b->d_G, b->kl_pad,
                           b->d_m2start, b->n_m2, b->d_TA, tap, b->jpad, b->d_TB, b->sep_kb, b->N, ncols, r->d_qhat,
                           r->d_qsum, r->d_tf, b->d_part, b->kl_pad, fwd_rows, b->n_m, b->d_gGp, b->kl_pad,
                           (int)b->nm_pad64);
    else   // the wide layout (odd-m classes, lambda dq != 0)
        hipLaunchKernelGGL(k_c3sep_bwd<kSepWideQB>, gbw, dim3(512), c3sep_lds_bytes(b->sep_kb), st, b->d_G,
                           b->kl_pad, b->d_m2start, b->n_m2, b->d_TA, tap, b->jpad, b->d_TB, b->sep_kb, b->N, ncols,
                           r->d_qhat, r->d_qsum, r->d_tf, b->d_part, b->kl_pad, fwd_rows, b->n_m, b->d_gGp,
                           b->kl_pad, (int)b->nm_pad64);
    TP_HIP(hipGetLastError());
    if (b->lam) {
        // dE/dG = the sum of the njc row blocks; dE/dAs[lt][kt] = sum over the
        // terms of lt of hw Y Y Y dE/dG[class][kt] (k_glam with the transposed tables)
        const int64_t n = (int64_t)b->n_m * b->kl_pad;
        hipLaunchKernelGGL(k_sum_blocks, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 8192)), dim3(256), 0, st,
                           b->d_gGp, b->njc, (int64_t)b->nm_pad64 * b->kl_pad, n, b->d_gG);
        hipLaunchKernelGGL(k_glam, dim3((unsigned)((b->kl_pad + b->lam_nt - 1) / b->lam_nt), (unsigned)b->n_lt),
                           dim3((unsigned)b->lam_nt), b->lamb_lds, st, b->d_gG, b->kl_pad, b->d_ktk, (int)b->klength,
                           b->d_lamY, b->K, b->d_lambyr, b->d_lambys, b->lamb_rows, b->d_lambent, b->d_lambcs,
                           b->d_gA, b->kl_pad, nullptr, nullptr);
    } else {
        const int n_rt = b->nlt_pad64 / G1_BM, n_ct = (int)(b->kl_pad / G1_BN);
        hipLaunchKernelGGL((k_gemm<G1_BM, G1_BN, G1_WGM, 2, false>), dim3((unsigned)(n_rt * n_ct)), dim3(256), 0, st,
                           b->d_HT4, (int64_t)b->nlt_pad64, b->d_gGp, b->kl_pad, (int)(b->njc * b->nm_pad64),
                           b->nlt_pad64, (int)b->kl_pad, n_rt, n_ct, b->d_gA, b->kl_pad, nullptr, nullptr, (int64_t)0,
                           nullptr);
    }
    TP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_triple_bwd, dim3((unsigned)b->n_lt), dim3(1024), b->bwd_lds, st, coef, b->n_c, b->K, b->d_ew,
                       b->d_eidx, b->d_gstart, b->d_ltinfo, b->gLw, b->gEmax, b->d_gA, b->kl_pad, (int)b->klength,
                       b->d_bwd_order, b->d_gpart);
    TP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_grad_reduce, dim3((unsigned)b->K, (unsigned)(b->L / 2 + 1)), dim3(64), 0, st, b->d_gpart, b->K,
                       b->gLw, b->n_c, b->d_coff, b->d_lstart, b->d_llist, b->d_grad);
    TP_HIP(hipGetLastError());
    return 0;
}

int tp_energy_grad(tp_ctx *ctx, const tp_basis *bc, const tp_ref *r, const double *coeff, int64_t ld, double *E_out,
                   double *grad_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !r || !E_out || !grad_out || !coeff) return tp_fail(TP_E_ARG, "null argument");
    tp_basis *b = const_cast<tp_basis *>(bc);
    if (b->ctx != ctx || r->b != b) return tp_fail(TP_E_STATE, "objects belong to another context/basis");
    if (ld < (int64_t)b->LMAX * b->LMAX) return tp_fail(TP_E_ARG, "coefficient vector shorter than LMAX^2");
    if (int rc = grad_setup(ctx, b)) return rc;
    if (int rc = upload_coeff(ctx, b, coeff, ld, b->K, b->d_coef, b->d_raw)) return rc;
    if (int rc = enqueue_eval(ctx, b, r, b->d_coef, nullptr, nullptr, 0, nullptr, b->d_out, mc_none(), true))
        return rc;
    if (int rc = enqueue_grad(ctx, b, r, b->d_coef)) return rc;
    std::vector<double2> g((size_t)b->K * b->n_c);
    TP_HIP(hipMemcpyAsync(ctx->h_pin, b->d_out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipMemcpyAsync(g.data(), b->d_grad, sizeof(double2) * g.size(), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    *E_out = ctx->h_pin[0];
    std::memset(grad_out, 0, sizeof(double) * 2 * (size_t)b->K * ld);
    for (int k = 0; k < b->K; ++k)
        for (int c = 0; c < b->n_c; ++c) {
            const size_t o = 2 * ((size_t)k * ld + b->sean0[c]);
            grad_out[o] = g[(size_t)k * b->n_c + c].x;
            grad_out[o + 1] = g[(size_t)k * b->n_c + c].y;
        }
    return 0;
}

// ---------------------------------------------------------------------------
// Photon-pair / photon-triplet histograms (tp_hist.hpp)
// ---------------------------------------------------------------------------
struct tp_hist {
    tp_ctx *ctx = nullptr;
    int N = 0, K2 = 0, K3 = 0;
    double dq = 0.0;
    int64_t nkt = 0;                 // k-triples k1 >= k2 >= k3 of 1..K3
    unsigned long long *d_n1 = nullptr, *d_n2 = nullptr, *d_n3 = nullptr;   // [K2], [N K2 K2], [nkt][2N][N]
    double *d_ph = nullptr;          // staging of the photons
    int64_t *d_off = nullptr;        // staging of the picture offsets
    int64_t ph_cap = 0, off_cap = 0;
    int64_t pictures = 0;
};

void tp_hist_destroy(tp_hist *h)
{
    if (!h) return;
    if (h->ctx) (void)hipSetDevice(h->ctx->device);
    void *ptrs[] = {h->d_n1, h->d_n2, h->d_n3, h->d_ph, h->d_off};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete h;
}

int tp_hist_create(tp_ctx *ctx, int N, int K2, int K3, double dq, tp_hist **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (int rc = check_ctx(ctx)) return rc;
    if (N < 1 || K2 < 1 || K3 < 1 || K3 > K2 || !(dq > 0.0)) return tp_fail(TP_E_ARG, "need N, K2 >= K3 >= 1 and dq > 0");
    tp_hist *h = new tp_hist();
    h->ctx = ctx;
    h->N = N;
    h->K2 = K2;
    h->K3 = K3;
    h->dq = dq;
    h->nkt = (int64_t)K3 * (K3 + 1) * (K3 + 2) / 6;
    int rc = 0;
    if ((rc = dev_alloc(&h->d_n1, (size_t)K2)) || (rc = dev_alloc(&h->d_n2, (size_t)N * K2 * K2)) ||
        (rc = dev_alloc(&h->d_n3, (size_t)h->nkt * 2 * N * N))) {
        tp_hist_destroy(h);
        return rc;
    }
    hipError_t e = hipMemset(h->d_n1, 0, sizeof(unsigned long long) * K2);
    if (e == hipSuccess) e = hipMemset(h->d_n2, 0, sizeof(unsigned long long) * (size_t)N * K2 * K2);
    if (e == hipSuccess) e = hipMemset(h->d_n3, 0, sizeof(unsigned long long) * (size_t)h->nkt * 2 * N * N);
    if (e != hipSuccess) {
        tp_hist_destroy(h);
        return tp_fail(TP_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    }
    *out = h;
    return 0;
}

int tp_hist_add(tp_hist *h, const double *photons, const int64_t *offsets, int npictures)
{
    if (!h || !offsets || npictures < 0) return tp_fail(TP_E_ARG, "bad argument");
    if (npictures == 0) return 0;
    if (int rc = check_ctx(h->ctx)) return rc;
    if (offsets[0] != 0) return tp_fail(TP_E_ARG, "offsets[0] must be 0");
    for (int i = 0; i < npictures; ++i) {
        const int64_t n = offsets[i + 1] - offsets[i];
        if (n < 0) return tp_fail(TP_E_ARG, "offsets must not decrease");
        if (n > kHistMaxPhotons) return tp_fail(TP_E_UNSUPPORTED, "a picture holds more than 512 photons");
    }
    const int64_t nph = offsets[npictures];
    if (nph > 0 && !photons) return tp_fail(TP_E_ARG, "null photons");
    if (3 * nph > h->ph_cap) {
        if (h->d_ph) TP_HIP(hipFree(h->d_ph));
        h->d_ph = nullptr;
        h->ph_cap = 0;
        if (int rc = dev_alloc(&h->d_ph, (size_t)std::max<int64_t>(3 * nph, 3))) return rc;
        h->ph_cap = std::max<int64_t>(3 * nph, 3);
    }
    if (npictures + 1 > h->off_cap) {
        if (h->d_off) TP_HIP(hipFree(h->d_off));
        h->d_off = nullptr;
        h->off_cap = 0;
        if (int rc = dev_alloc(&h->d_off, (size_t)npictures + 1)) return rc;
        h->off_cap = npictures + 1;
    }
    tp_ctx *ctx = h->ctx;
    hipStream_t st = ctx->stream;
    if (nph > 0) TP_HIP(hipMemcpyAsync(h->d_ph, photons, sizeof(double) * 3 * nph, hipMemcpyHostToDevice, st));
    TP_HIP(hipMemcpyAsync(h->d_off, offsets, sizeof(int64_t) * (npictures + 1), hipMemcpyHostToDevice, st));
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[0], st));
    hipLaunchKernelGGL(k_hist_pictures, dim3((unsigned)npictures), dim3(kHistThreads), 0, st, h->d_ph, h->d_off,
                       (const int *)nullptr, 0, h->dq, h->N, h->K2, h->K3, h->d_n1, h->d_n2, h->d_n3);
    TP_HIP(hipGetLastError());
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[1], st));
    TP_HIP(hipStreamSynchronize(st));
    if (ctx->timing) {
        float t = 0;
        if (hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->last_ms[0] = t;
        for (int i = 1; i < 4; ++i) ctx->last_ms[i] = 0.0;
    }
    h->pictures += npictures;
    return 0;
}

int tp_hist_read(tp_hist *h, double *c1, double *c2, double *c3, uint64_t *n1_out, uint64_t *n2_out,
                 uint64_t *n3_out)
{
    if (!h) return tp_fail(TP_E_ARG, "null histogram");
    if (int rc = check_ctx(h->ctx)) return rc;
    const int N = h->N, K2 = h->K2, K3 = h->K3, R = 2 * N * N;
    std::vector<unsigned long long> n1(K2), n2((size_t)N * K2 * K2), n3((size_t)h->nkt * R);
    TP_HIP(hipMemcpy(n1.data(), h->d_n1, sizeof(unsigned long long) * n1.size(), hipMemcpyDeviceToHost));
    TP_HIP(hipMemcpy(n2.data(), h->d_n2, sizeof(unsigned long long) * n2.size(), hipMemcpyDeviceToHost));
    TP_HIP(hipMemcpy(n3.data(), h->d_n3, sizeof(unsigned long long) * n3.size(), hipMemcpyDeviceToHost));
    for (int k1 = 1; k1 <= K2; ++k1) {
        if (c1) c1[k1 - 1] = (double)n1[k1 - 1] * (1.0 / k1);
        if (n1_out) n1_out[k1 - 1] = n1[k1 - 1];
        for (int k2 = 1; k2 <= K2; ++k2) {
            const double w = 1.0 / ((k1 != k2 ? 2.0 : 1.0) * k1 * k2);
            for (int a = 0; a < N; ++a) {
                const size_t o = a + (size_t)N * ((k2 - 1) + (size_t)K2 * (k1 - 1));
                if (c2) c2[o] = (double)n2[o] * w;
                if (n2_out) n2_out[o] = n2[o];
            }
        }
    }
    const size_t full = (size_t)R * K3 * K3 * K3;
    if (c3) std::memset(c3, 0, sizeof(double) * full);
    if (n3_out) std::memset(n3_out, 0, sizeof(uint64_t) * full);
    int64_t kt = 0;
    for (int k1 = 1; k1 <= K3; ++k1)
        for (int k2 = 1; k2 <= k1; ++k2)
            for (int k3 = 1; k3 <= k2; ++k3, ++kt) {
                const double tf = (k1 == k2 && k2 == k3) ? 1.0 : ((k1 == k2 || k2 == k3 || k1 == k3) ? 3.0 : 6.0);
                const double w = 1.0 / (tf * k1 * k2 * k3);
                const size_t o = (size_t)R * ((k3 - 1) + (size_t)K3 * ((k2 - 1) + (size_t)K3 * (k1 - 1)));
                for (int r = 0; r < R; ++r) {
                    const unsigned long long c = n3[(size_t)kt * R + r];
                    if (c3) c3[o + r] = (double)c * w;
                    if (n3_out) n3_out[o + r] = c;
                }
            }
    return 0;
}

// twoPhotons with ewc = dq lambda / (4 pi) (0: the flat detector)
static int two_photons(tp_ctx *ctx, const tp_basis *bc, const double *coeff, int64_t ld, int K2, int minimal,
                       int normalize, double ewc, double *c2_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !c2_out) return tp_fail(TP_E_ARG, "null argument");
    tp_basis *b = const_cast<tp_basis *>(bc);
    if (b->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
    if (K2 < 1) return tp_fail(TP_E_ARG, "need K2 >= 1");
    if (!(std::fabs(K2 * ewc) <= 1.0)) return tp_fail(TP_E_ARG, "alpha_star undefined: need |K2 dq lambda| <= 4 pi");
    const int N = b->N, nl = b->L / b->lstep + 1;
    const size_t lds = two_photons_lds_bytes(b->n_c, N, nl);
    if (lds > 64 * 1024) return tp_fail(TP_E_UNSUPPORTED, "two-photon tables exceed 64 KiB of LDS (N or L too large)");
    const int64_t len = (int64_t)N * K2 * K2;
    if (K2 > b->c2_shells) {
        if (b->d_c2coef) TP_HIP(hipFree(b->d_c2coef));
        if (b->d_c2raw) TP_HIP(hipFree(b->d_c2raw));
        b->d_c2coef = nullptr;
        b->d_c2raw = nullptr;
        b->c2_shells = 0;
        if (int rc = dev_alloc(&b->d_c2coef, (size_t)K2 * b->n_c)) return rc;
        if (int rc = dev_alloc(&b->d_c2raw, (size_t)K2 * b->LMAX * b->LMAX)) return rc;
        b->c2_shells = K2;
    }
    if (len > b->c2_len) {
        if (b->d_c2) TP_HIP(hipFree(b->d_c2));
        b->d_c2 = nullptr;
        b->c2_len = 0;
        if (int rc = dev_alloc(&b->d_c2, (size_t)len)) return rc;
        b->c2_len = len;
    }
    // upload_coeff stages through d_c2raw only for ld == LMAX^2 and K2 <= K
    // shells; otherwise it allocates its own staging buffer.
    if (int rc = upload_coeff(ctx, b, coeff, ld, K2, b->d_c2coef, K2 <= b->K ? b->d_c2raw : nullptr)) return rc;
    hipStream_t st = ctx->stream;
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[0], st));
    hipLaunchKernelGGL(k_two_photons, dim3((unsigned)K2, (unsigned)((K2 + kC2Chunk - 1) / kC2Chunk)), dim3(256), lds, st,
                       b->d_c2coef, b->n_c, b->d_coff, b->L,
                       b->lstep, N, K2, minimal ? 1 : 0, normalize ? 1.0 : 1.0 / (4.0 * tph::kPi), ewc, b->d_c2);
    TP_HIP(hipGetLastError());
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[1], st));
    TP_HIP(hipMemcpyAsync(c2_out, b->d_c2, sizeof(double) * (size_t)len, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    if (ctx->timing) {
        float t = 0;
        if (hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->last_ms[0] = t;
        for (int i = 1; i < 4; ++i) ctx->last_ms[i] = 0.0;
    }
    if (normalize) {   // c / sum(abs, c)
        double s = 0.0;
        for (int64_t i = 0; i < len; ++i) s += std::fabs(c2_out[i]);
        for (int64_t i = 0; i < len; ++i) c2_out[i] /= s;
    }
    return 0;
}

int tp_two_photons(tp_ctx *ctx, const tp_basis *bc, const double *coeff, int64_t ld, int K2, int minimal,
                   int normalize, double *c2_out)
{
    if (bc && bc->lam)
        return tp_fail(TP_E_UNSUPPORTED, "the basis has lambda dq != 0: use tp_two_photons_ewald with the volume's dq");
    return two_photons(ctx, bc, coeff, ld, K2, minimal, normalize, 0.0, c2_out);
}

int tp_two_photons_ewald(tp_ctx *ctx, const tp_basis *bc, const double *coeff, int64_t ld, int K2, int minimal,
                         int normalize, double dq, double *c2_out)
{
    if (!bc) return tp_fail(TP_E_ARG, "null basis");
    return two_photons(ctx, bc, coeff, ld, K2, minimal, normalize, dq * bc->lambda / (4 * tph::kPi), c2_out);
}

int tp_retrieve_solution(const double *c2, int N, int K2dim, int k2lo, int k2hi, int L, int LMAX, double *coeff_out,
                         int64_t ld, int *flipped, double *negativity)
{
    if (!c2 || !coeff_out) return tp_fail(TP_E_ARG, "null argument");
    if (N < 1 || K2dim < 1 || LMAX < 1) return tp_fail(TP_E_ARG, "need N, K2, LMAX >= 1");
    const std::string err = tpc2::retrieve_solution(c2, N, K2dim, k2lo, k2hi, L, LMAX,
                                                    reinterpret_cast<std::complex<double> *>(coeff_out), ld, flipped,
                                                    negativity);
    if (!err.empty()) return tp_fail(TP_E_ARG, err);
    return 0;
}

int tp_retrieve_solution_ewald(const double *c2, int N, int K2dim, int k2lo, int k2hi, int L, int LMAX, double qmax,
                               double lambda, double *coeff_out, int64_t ld, int *flipped, double *negativity)
{
    if (!c2 || !coeff_out) return tp_fail(TP_E_ARG, "null argument");
    if (N < 1 || K2dim < 1 || LMAX < 1 || k2hi < 1) return tp_fail(TP_E_ARG, "need N, K2, LMAX, k2hi >= 1");
    // the structure is SphericalHarmonicsVolume(LMAX, k2hi, qmax): dq = qmax / k2hi (correlations.jl:269-271)
    const double mdq = qmax / k2hi;
    if (!(std::fabs(k2hi * mdq * lambda / (4 * tph::kPi)) <= 1.0))
        return tp_fail(TP_E_ARG, "alpha_star undefined: need |qmax lambda| <= 4 pi");
    const std::string err = tpc2::retrieve_solution(c2, N, K2dim, k2lo, k2hi, L, LMAX,
                                                    reinterpret_cast<std::complex<double> *>(coeff_out), ld, flipped,
                                                    negativity, lambda, mdq);
    if (!err.empty()) return tp_fail(TP_E_ARG, err);
    return 0;
}

// ---------------------------------------------------------------------------
// rotational fit (fitStructures_*, data_processing.jl:26-79)
// ---------------------------------------------------------------------------
struct tp_fit {
    tp_ctx *ctx = nullptr;
    tp_basis *b = nullptr;
    int K = 0, klo = 0, nk = 0;       // shells 1..K rotated, similarity over klo+1..K
    double *d_P = nullptr;            // [LMAX][LMAX][2 LMAX] sphPlm(l, m, theta_j) at ((m LMAX + l) 2 LMAX + j)
    int *d_sean = nullptr;            // [LMAX^2] seanindex - 1 of (l, m) at l^2 + l + m
    double2 *d_gref = nullptr;        // [nk][2 LMAX - 1][2 LMAX] reference g_j(m)
    double *d_refsum = nullptr;       // [nk][2] sum Re h, sum (Re h)^2
    double2 *d_vol = nullptr;         // [K][LMAX^2] the structure (full layout)
    double2 *d_volc = nullptr;        // [K][n_c] the structure (compact)
    bool have_vol = false;
    int cap = 0;                      // rotations per chunk the buffers hold
    double2 *d_rot = nullptr;         // [cap][K][n_c]
    double *d_quat = nullptr;         // [cap][L/2][4]
    EvalDesc *d_ev = nullptr;         // [cap]
    double *d_cor = nullptr;          // [cap][nk]
};

void tp_fit_destroy(tp_fit *f)
{
    if (!f) return;
    if (f->ctx) (void)hipSetDevice(f->ctx->device);
    void *ptrs[] = {f->d_P, f->d_sean, f->d_gref, f->d_refsum, f->d_vol, f->d_volc, f->d_rot, f->d_quat, f->d_ev,
                    f->d_cor};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete f;
}

// Copy `nshell` shells of host coefficients (stride ld complex) to a device
// [nshell][LMAX^2] array.
static int upload_full(tp_ctx *ctx, const double *coeff, int64_t ld, int nshell, int LMAX, double2 *dst)
{
    const size_t w = sizeof(double2) * (size_t)LMAX * LMAX;
    TP_HIP(hipMemcpy2DAsync(dst, w, coeff, sizeof(double2) * (size_t)ld, w, (size_t)nshell, hipMemcpyHostToDevice,
                            ctx->stream));
    return 0;
}

int tp_fit_create(tp_ctx *ctx, const tp_basis *bc, const double *reference, int64_t ld, int k_lo, int k_hi,
                  tp_fit **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !reference) return tp_fail(TP_E_ARG, "null argument");
    if (bc->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
    tp_basis *b = const_cast<tp_basis *>(bc);
    const int LMAX = b->LMAX;
    if (b->L < 2 || b->L > kRotMaxL) return tp_fail(TP_E_UNSUPPORTED, "rotational fit needs 2 <= L <= 22");
    if (LMAX > kFitMaxLMAX) return tp_fail(TP_E_UNSUPPORTED, "rotational fit needs LMAX <= 32");
    if (k_lo < 1 || k_hi < k_lo) return tp_fail(TP_E_ARG, "need 1 <= k_lo <= k_hi");
    if (ld < (int64_t)LMAX * LMAX) return tp_fail(TP_E_ARG, "ld shorter than LMAX^2");
    tp_fit *f = new tp_fit();
    f->ctx = ctx;
    f->b = b;
    f->K = k_hi;
    f->klo = k_lo - 1;
    f->nk = k_hi - k_lo + 1;
    const int S = 2 * LMAX, nmm = 2 * LMAX - 1;
    std::vector<double> P((size_t)S * LMAX * LMAX, 0.0);
    for (int j = 0; j < S; ++j) {
        const double th = tph::kPi * (2 * j + 1) / (4.0 * LMAX);
        for (int m = 0; m < LMAX; ++m)
            for (int l = m; l < LMAX; ++l) P[((size_t)m * LMAX + l) * S + j] = tph::sphPlm(l, m, th);
    }
    std::vector<int> sean((size_t)LMAX * LMAX);
    for (int l = 0; l < LMAX; ++l)
        for (int m = -l; m <= l; ++m) sean[l * l + l + m] = (int)(tph::seanindex(l, m, LMAX) - 1);
    int rc = 0;
    if ((rc = dev_alloc(&f->d_P, P.size())) || (rc = dev_alloc(&f->d_sean, sean.size())) ||
        (rc = dev_alloc(&f->d_gref, (size_t)f->nk * S * nmm)) || (rc = dev_alloc(&f->d_refsum, (size_t)2 * f->nk)) ||
        (rc = dev_alloc(&f->d_vol, (size_t)k_hi * LMAX * LMAX)) || (rc = dev_alloc(&f->d_volc, (size_t)k_hi * b->n_c))) {
        tp_fit_destroy(f);
        return rc;
    }
    hipStream_t st = ctx->stream;
    hipError_t e = hipMemcpyAsync(f->d_P, P.data(), sizeof(double) * P.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(f->d_sean, sean.data(), sizeof(int) * sean.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        const size_t w = sizeof(double2) * (size_t)LMAX * LMAX;
        e = hipMemcpy2DAsync(f->d_vol, w, reference, sizeof(double2) * (size_t)ld, w, (size_t)k_hi,
                             hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fit_similarity, dim3((unsigned)f->nk, 1u), dim3(kFitThreads), 0, st, f->d_vol,
                           (int64_t)LMAX * LMAX, (const double2 *)nullptr, b->n_c, b->d_coff, b->L, LMAX, f->K, f->klo,
                           f->nk, f->d_sean, f->d_P, f->d_gref, f->d_refsum, (double *)nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        tp_fit_destroy(f);
        return tp_fail(TP_E_HIP, std::string("tp_fit_create: ") + hipGetErrorString(e));
    }
    *out = f;
    return 0;
}

int tp_fit_set_volume(tp_fit *f, const double *coeff, int64_t ld)
{
    if (!f || !coeff) return tp_fail(TP_E_ARG, "null argument");
    if (int rc = check_ctx(f->ctx)) return rc;
    const int LMAX = f->b->LMAX;
    if (ld < (int64_t)LMAX * LMAX) return tp_fail(TP_E_ARG, "ld shorter than LMAX^2");
    if (int rc = upload_full(f->ctx, coeff, ld, f->K, LMAX, f->d_vol)) return rc;
    if (int rc = upload_coeff(f->ctx, f->b, coeff, ld, f->K, f->d_volc, nullptr)) return rc;
    TP_HIP(hipStreamSynchronize(f->ctx->stream));
    f->have_vol = true;
    return 0;
}

// Unit quaternion (w, x, y, z) of R = Rz(phi) Ry(-theta) Rz(gamma): the
// rotation rotationMatrix(theta, phi, gamma, l) applies to the coefficients
// (spherical_harmonics.jl:589-595; the rotated structure is f(R^-1 r)).
static void euler_quat(double theta, double phi, double gamma, double *q)
{
    const double a[4] = {std::cos(0.5 * phi), 0.0, 0.0, std::sin(0.5 * phi)};
    const double bq[4] = {std::cos(-0.5 * theta), 0.0, std::sin(-0.5 * theta), 0.0};
    const double c[4] = {std::cos(0.5 * gamma), 0.0, 0.0, std::sin(0.5 * gamma)};
    auto mul = [](const double *u, const double *v, double *o) {
        o[0] = u[0] * v[0] - u[1] * v[1] - u[2] * v[2] - u[3] * v[3];
        o[1] = u[0] * v[1] + u[1] * v[0] + u[2] * v[3] - u[3] * v[2];
        o[2] = u[0] * v[2] - u[1] * v[3] + u[2] * v[0] + u[3] * v[1];
        o[3] = u[0] * v[3] + u[1] * v[2] - u[2] * v[1] + u[3] * v[0];
    };
    double t[4];
    mul(a, bq, t);
    mul(t, c, q);
}

int tp_fit_similarity(tp_fit *f, int n, const double *euler, double *sc_out)
{
    if (!f || (n > 0 && (!euler || !sc_out)) || n < 0) return tp_fail(TP_E_ARG, "bad argument");
    if (n == 0) return 0;
    if (int rc = check_ctx(f->ctx)) return rc;
    if (!f->have_vol) return tp_fail(TP_E_STATE, "tp_fit_set_volume first");
    tp_ctx *ctx = f->ctx;
    tp_basis *b = f->b;
    hipStream_t st = ctx->stream;
    const int nl = b->L / 2, chunk = std::min(n, 2048);
    if (chunk > f->cap) {
        void *ptrs[] = {f->d_rot, f->d_quat, f->d_ev, f->d_cor};
        for (void *p : ptrs)
            if (p) TP_HIP(hipFree(p));
        f->d_rot = nullptr; f->d_quat = nullptr; f->d_ev = nullptr; f->d_cor = nullptr;
        f->cap = 0;
        int rc = 0;
        if ((rc = dev_alloc(&f->d_rot, (size_t)chunk * f->K * b->n_c)) ||
            (rc = dev_alloc(&f->d_quat, (size_t)chunk * nl * 4)) || (rc = dev_alloc(&f->d_ev, (size_t)chunk)) ||
            (rc = dev_alloc(&f->d_cor, (size_t)chunk * f->nk)))
            return rc;
        f->cap = chunk;
    }
    std::vector<double> quat((size_t)chunk * nl * 4);
    std::vector<EvalDesc> ev(chunk);
    std::vector<double> cor((size_t)chunk * f->nk);
    float t_rot = 0.f, t_sim = 0.f;
    for (int r0 = 0; r0 < n; r0 += chunk) {
        const int c = std::min(chunk, n - r0);
        for (int i = 0; i < c; ++i) {
            double q[4];
            euler_quat(euler[3 * (r0 + i)], euler[3 * (r0 + i) + 1], euler[3 * (r0 + i) + 2], q);
            for (int li = 0; li < nl; ++li) std::memcpy(&quat[((size_t)i * nl + li) * 4], q, sizeof(q));
            EvalDesc &d = ev[i];
            std::memset(&d, 0, sizeof(d));
            d.src = f->d_volc;
            d.coef = f->d_rot + (size_t)i * f->K * b->n_c;
            d.K = f->K;
        }
        TP_HIP(hipMemcpyAsync(f->d_quat, quat.data(), sizeof(double) * (size_t)c * nl * 4, hipMemcpyHostToDevice, st));
        TP_HIP(hipMemcpyAsync(f->d_ev, ev.data(), sizeof(EvalDesc) * (size_t)c, hipMemcpyHostToDevice, st));
        if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[0], st));
        hipLaunchKernelGGL(k_rotate_batch, dim3((unsigned)nl, (unsigned)c), dim3(256), 0, st, f->d_ev, f->d_quat, nl,
                           b->n_c, (const GroupState *)nullptr, TP_MOVE_QUAT, 0);
        TP_HIP(hipGetLastError());
        if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[1], st));
        hipLaunchKernelGGL(k_fit_similarity, dim3((unsigned)f->nk, (unsigned)c), dim3(kFitThreads), 0, st, f->d_vol,
                           (int64_t)b->LMAX * b->LMAX, (const double2 *)f->d_rot, b->n_c, b->d_coff, b->L, b->LMAX,
                           f->K, f->klo, f->nk, f->d_sean, f->d_P, f->d_gref, f->d_refsum, f->d_cor);
        TP_HIP(hipGetLastError());
        if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[2], st));
        TP_HIP(hipMemcpyAsync(cor.data(), f->d_cor, sizeof(double) * (size_t)c * f->nk, hipMemcpyDeviceToHost, st));
        TP_HIP(hipStreamSynchronize(st));
        if (ctx->timing) {
            float a = 0.f, bb = 0.f;
            if (hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]) == hipSuccess) t_rot += a;
            if (hipEventElapsedTime(&bb, ctx->ev[1], ctx->ev[2]) == hipSuccess) t_sim += bb;
        }
        for (int i = 0; i < c; ++i) {
            double s = 0.0;
            for (int kk = 0; kk < f->nk; ++kk) s += cor[(size_t)i * f->nk + kk];
            sc_out[r0 + i] = s / f->nk;
        }
    }
    if (ctx->timing) {
        ctx->last_ms[0] = t_rot;
        ctx->last_ms[1] = t_sim;
        ctx->last_ms[2] = ctx->last_ms[3] = 0.0;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// SH transforms, spherical Bessel transform and sRAAR (phases.jl:31-95)
// ---------------------------------------------------------------------------
struct PhTables {
    int LMAX = 0, K = 0;
    double *d_P = nullptr;    // [LMAX][LMAX][2 LMAX] sphPlm(l, m, theta_j) at (m LMAX + l) 2 LMAX + j
    double *d_W = nullptr;    // [2 LMAX] (2 pi / 2 LMAX) w_j, Driscoll-Healy weights
    double2 *d_tw = nullptr;  // [2 LMAX] e^{2 pi i r / 2 LMAX}
    double *d_Hf = nullptr;   // [LMAX][K][K] forward spherical Bessel matrices
    double *d_Hb = nullptr;   // [LMAX][K][K] backward
    double2 *d_G = nullptr;   // [K][2 LMAX][2 LMAX - 1] phi sums of the analysis
    ~PhTables()
    {
        void *ptrs[] = {d_P, d_W, d_tw, d_Hf, d_Hb, d_G};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
};

// frdr weights of the reference's Hankel quadrature (spherical_harmonics.jl:296-307):
// w_j = pi/4 res_j^1.5 with res from the sample positions x_1..x_K
static std::vector<double> ph_frdr(const std::vector<double> &x)
{
    const int n = (int)x.size();
    std::vector<double> res(n);
    res[0] = std::sqrt(0.5) * (x[0] + x[1]) * (x[0] + x[1]);
    for (int j = 1; j < n - 1; ++j) res[j] = (x[j + 1] - x[j - 1]) * (x[j - 1] + 2 * x[j] + x[j + 1]);
    res[n - 1] = 4 * x[n - 1] * x[n - 1] - (x[n - 2] + x[n - 1]) * (x[n - 2] + x[n - 1]);
    for (int j = 0; j < n; ++j) res[j] = tph::kPi / 4 * std::pow(res[j], 1.5);
    return res;
}

static int ph_tables(tp_ctx *ctx, int LMAX, int K, bool bessel, PhTables &t)
{
    if (LMAX < 1 || LMAX > kPhMaxLMAX) return tp_fail(TP_E_UNSUPPORTED, "need 1 <= LMAX <= 48");
    if (K < 1 || (bessel && K < 2)) return tp_fail(TP_E_ARG, "need K >= 1 (>= 2 for the Bessel transform)");
    t.LMAX = LMAX;
    t.K = K;
    const int S = 2 * LMAX;
    std::vector<double> P((size_t)LMAX * LMAX * S, 0.0), W(S);
    std::vector<double2> tw(S);
    for (int j = 0; j < S; ++j) {
        const double th = tph::kPi * (2 * j + 1) / (4.0 * LMAX);
        for (int m = 0; m < LMAX; ++m)
            for (int l = m; l < LMAX; ++l) P[((size_t)m * LMAX + l) * S + j] = tph::sphPlm(l, m, th);
        // Driscoll-Healy: w_j = (2/B) sin(theta_j) sum_{k<B} sin((2k+1) theta_j) / (2k+1)
        double sk = 0.0;
        for (int k = 0; k < LMAX; ++k) sk += std::sin((2 * k + 1) * th) / (2 * k + 1);
        W[j] = (2.0 * tph::kPi / S) * (2.0 / LMAX) * std::sin(th) * sk;
        tw[j] = make_double2(std::cos(2.0 * tph::kPi * j / S), std::sin(2.0 * tph::kPi * j / S));
    }
    int rc = 0;
    if ((rc = dev_alloc(&t.d_P, P.size())) || (rc = dev_alloc(&t.d_W, W.size())) || (rc = dev_alloc(&t.d_tw, tw.size())) ||
        (rc = dev_alloc(&t.d_G, (size_t)K * S * (2 * LMAX - 1))))
        return rc;
    TP_HIP(hipMemcpyAsync(t.d_P, P.data(), sizeof(double) * P.size(), hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipMemcpyAsync(t.d_W, W.data(), sizeof(double) * W.size(), hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipMemcpyAsync(t.d_tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice, ctx->stream));
    if (bessel) {
        // sphBesselTransform (spherical_harmonics.jl:310-377): I[k][r] = J_{l+1/2}(pi k r / K);
        // forward out[k] = i^l sqrt(1/k) sum_r I[k][r] frdr(r)_r in[r], r = 1..K;
        // backward out[r] = (-i)^l sqrt(1/r) sum_k I[k][r] frdr(q)_k in[k] / (2 pi)^2, q = pi k / K
        std::vector<double> xr(K), xq(K);
        for (int i = 0; i < K; ++i) {
            xr[i] = i + 1.0;
            xq[i] = tph::kPi / K * (i + 1.0);
        }
        const std::vector<double> wr = ph_frdr(xr), wq = ph_frdr(xq);
        std::vector<double> Hf((size_t)LMAX * K * K), Hb((size_t)LMAX * K * K), J((size_t)LMAX);
        for (int k = 0; k < K; ++k)
            for (int r = 0; r < K; ++r) {
                tph::bessel_half_all(xq[k] * xr[r], LMAX, J.data());   // every order at this argument
                for (int l = 0; l < LMAX; ++l) {
                    const double I = J[(size_t)l];
                    Hf[((size_t)l * K + k) * K + r] = std::sqrt(1.0 / (k + 1)) * I * wr[r];
                    Hb[((size_t)l * K + r) * K + k] = std::sqrt(1.0 / (r + 1)) * I * wq[k] / (4.0 * tph::kPi * tph::kPi);
                }
            }
        if ((rc = dev_alloc(&t.d_Hf, Hf.size())) || (rc = dev_alloc(&t.d_Hb, Hb.size()))) return rc;
        TP_HIP(hipMemcpyAsync(t.d_Hf, Hf.data(), sizeof(double) * Hf.size(), hipMemcpyHostToDevice, ctx->stream));
        TP_HIP(hipMemcpyAsync(t.d_Hb, Hb.data(), sizeof(double) * Hb.size(), hipMemcpyHostToDevice, ctx->stream));
    }
    TP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

// Tables of (LMAX, K) from the context's cache, built on a miss (a table with
// the Bessel matrices also serves calls without them).  The cache keeps the
// 8 most recent sizes; the analysis scratch d_G in it is shared by the calls
// of one context, which never overlap.
static int ph_get(tp_ctx *ctx, int LMAX, int K, bool bessel, const PhTables **out)
{
    auto &c = ctx->ph_cache;
    for (size_t i = 0; i < c.size(); ++i)
        if (c[i]->LMAX == LMAX && c[i]->K == K && (!bessel || c[i]->d_Hf)) {
            *out = c[i];
            return 0;
        }
    PhTables *t = new PhTables();
    if (int rc = ph_tables(ctx, LMAX, K, bessel, *t)) {
        delete t;
        return rc;
    }
    if (c.size() >= 8) {
        delete c.front();
        c.erase(c.begin());
    }
    c.push_back(t);
    *out = t;
    return 0;
}

static void ph_cache_free(tp_ctx *ctx)
{
    for (PhTables *t : ctx->ph_cache) delete t;
    ctx->ph_cache.clear();
}

// host coefficients (seanindex layout, stride ld) <-> packed (l^2 + l + m)
static std::vector<double2> ph_pack(const double *coeff, int64_t ld, int K, int LMAX)
{
    std::vector<double2> out((size_t)K * LMAX * LMAX);
    for (int k = 0; k < K; ++k)
        for (int l = 0; l < LMAX; ++l)
            for (int m = -l; m <= l; ++m) {
                const int64_t s = tph::seanindex(l, m, LMAX) - 1;
                out[(size_t)k * LMAX * LMAX + l * l + l + m] =
                    make_double2(coeff[2 * ((size_t)k * ld + s)], coeff[2 * ((size_t)k * ld + s) + 1]);
            }
    return out;
}

static void ph_unpack(const std::vector<double2> &in, int K, int LMAX, double *coeff, int64_t ld)
{
    std::memset(coeff, 0, sizeof(double) * 2 * (size_t)K * ld);
    for (int k = 0; k < K; ++k)
        for (int l = 0; l < LMAX; ++l)
            for (int m = -l; m <= l; ++m) {
                const int64_t s = tph::seanindex(l, m, LMAX) - 1;
                const double2 v = in[(size_t)k * LMAX * LMAX + l * l + l + m];
                coeff[2 * ((size_t)k * ld + s)] = v.x;
                coeff[2 * ((size_t)k * ld + s) + 1] = v.y;
            }
}

static void ph_synth(tp_ctx *ctx, const PhTables &t, const double2 *c, double2 *surf)
{
    const unsigned nch = (unsigned)((2 * t.LMAX + kPhRings - 1) / kPhRings);
    hipLaunchKernelGGL(k_sh_synth, dim3((unsigned)t.K, nch), dim3(kPhThreads), sh_synth_lds_bytes(t.LMAX), ctx->stream,
                       c, t.LMAX, t.d_P, t.d_tw, surf);
}

static void ph_analysis(tp_ctx *ctx, const PhTables &t, const double2 *surf, double2 *c)
{
    const unsigned nch = (unsigned)((2 * t.LMAX + kPhRings - 1) / kPhRings);
    hipLaunchKernelGGL(k_sh_rings, dim3((unsigned)t.K, nch), dim3(kPhThreads), sh_rings_lds_bytes(t.LMAX), ctx->stream,
                       surf, t.LMAX, t.d_tw, t.d_G);
    const int n = t.K * t.LMAX * t.LMAX;
    hipLaunchKernelGGL(k_sh_analysis, dim3((unsigned)((n + kPhThreads - 1) / kPhThreads)), dim3(kPhThreads), 0,
                       ctx->stream, t.d_G, t.LMAX, t.K, t.d_P, t.d_W, c);
}

static void ph_bessel(tp_ctx *ctx, const PhTables &t, const double2 *in, int forward, double2 *out)
{
    const int n = t.K * t.LMAX * t.LMAX;
    hipLaunchKernelGGL(k_sph_bessel, dim3((unsigned)((n + kPhThreads - 1) / kPhThreads)), dim3(kPhThreads), 0,
                       ctx->stream, in, t.LMAX, t.K, forward ? t.d_Hf : t.d_Hb, forward, out);
}

extern "C++" {
template <typename T>
struct DevBuf {   // device buffer freed with its scope
    T *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};
}

int tp_sh_surface(tp_ctx *ctx, int LMAX, int K, const double *coeff, int64_t ld, double *surf)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!coeff || !surf || ld < (int64_t)LMAX * LMAX) return tp_fail(TP_E_ARG, "bad argument");
    const PhTables *tcache = nullptr;
    if (int rc = ph_get(ctx, LMAX, K, false, &tcache)) return rc;
    const PhTables &t = *tcache;
    const size_t NP = (size_t)4 * LMAX * LMAX;
    DevBuf<double2> c, f;
    if (int rc = dev_alloc(&c.p, (size_t)K * LMAX * LMAX)) return rc;
    if (int rc = dev_alloc(&f.p, (size_t)K * NP)) return rc;
    const std::vector<double2> h = ph_pack(coeff, ld, K, LMAX);
    TP_HIP(hipMemcpyAsync(c.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    ph_synth(ctx, t, c.p, f.p);
    TP_HIP(hipGetLastError());
    TP_HIP(hipMemcpyAsync(surf, f.p, sizeof(double2) * K * NP, hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tp_sh_coeff(tp_ctx *ctx, int LMAX, int K, const double *surf, double *coeff, int64_t ld)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!coeff || !surf || ld < (int64_t)LMAX * LMAX) return tp_fail(TP_E_ARG, "bad argument");
    const PhTables *tcache = nullptr;
    if (int rc = ph_get(ctx, LMAX, K, false, &tcache)) return rc;
    const PhTables &t = *tcache;
    const size_t NP = (size_t)4 * LMAX * LMAX;
    DevBuf<double2> c, f;
    if (int rc = dev_alloc(&c.p, (size_t)K * LMAX * LMAX)) return rc;
    if (int rc = dev_alloc(&f.p, (size_t)K * NP)) return rc;
    TP_HIP(hipMemcpyAsync(f.p, surf, sizeof(double2) * K * NP, hipMemcpyHostToDevice, ctx->stream));
    ph_analysis(ctx, t, f.p, c.p);
    TP_HIP(hipGetLastError());
    std::vector<double2> h((size_t)K * LMAX * LMAX);
    TP_HIP(hipMemcpyAsync(h.data(), c.p, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    ph_unpack(h, K, LMAX, coeff, ld);
    return 0;
}

int tp_sph_bessel(tp_ctx *ctx, int LMAX, int K, int forward, const double *coeff, int64_t ld, double *out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!coeff || !out || ld < (int64_t)LMAX * LMAX) return tp_fail(TP_E_ARG, "bad argument");
    const PhTables *tcache = nullptr;
    if (int rc = ph_get(ctx, LMAX, K, true, &tcache)) return rc;
    const PhTables &t = *tcache;
    DevBuf<double2> a, b;
    if (int rc = dev_alloc(&a.p, (size_t)K * LMAX * LMAX)) return rc;
    if (int rc = dev_alloc(&b.p, (size_t)K * LMAX * LMAX)) return rc;
    const std::vector<double2> h = ph_pack(coeff, ld, K, LMAX);
    TP_HIP(hipMemcpyAsync(a.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    ph_bessel(ctx, t, a.p, forward, b.p);
    TP_HIP(hipGetLastError());
    std::vector<double2> r(h.size());
    TP_HIP(hipMemcpyAsync(r.data(), b.p, sizeof(double2) * r.size(), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    ph_unpack(r, K, LMAX, out, ld);
    return 0;
}

// Atoms (x, y, z, R, B) -> the density kernel's rows, sorted by |a|, and per
// shell r_k = k rmax / K the range of atoms with ||a| - r_k| < sqrt 5 (plus
// a margin for rounding): rows followed by the K ranges.  R, B > 0.
static int str_atoms(const double *atoms, int natoms, int K, double rmax, std::vector<StrAtom> &out,
                     std::vector<int2> &range)
{
    std::vector<double> rad((size_t)natoms);
    std::vector<int> ord((size_t)natoms);
    for (int a = 0; a < natoms; ++a) {
        const double *r = atoms + (size_t)5 * a;
        if (!(r[3] > 0.0) || !(r[4] > 0.0)) return tp_fail(TP_E_ARG, "atom form factors need R > 0 and B > 0");
        rad[(size_t)a] = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        ord[(size_t)a] = a;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return rad[(size_t)x] < rad[(size_t)y]; });
    out.resize((size_t)natoms);
    std::vector<double> srad((size_t)natoms);
    for (int i = 0; i < natoms; ++i) {
        const double *r = atoms + (size_t)5 * ord[(size_t)i];
        const double R = r[3], s = R * std::sqrt(2.0 * tph::kPi);
        out[(size_t)i] = StrAtom{r[0], r[1], r[2], r[4] / (s * s * s), 2.0 * R * R};
        srad[(size_t)i] = rad[(size_t)ord[(size_t)i]];
    }
    const double reach = std::sqrt(5.0) + 1e-6;
    range.resize((size_t)K);
    for (int k = 0; k < K; ++k) {
        const double rk = (k + 1) * (rmax / K);
        const int lo = (int)(std::lower_bound(srad.begin(), srad.end(), rk - reach) - srad.begin());
        const int hi = (int)(std::upper_bound(srad.begin(), srad.end(), rk + reach) - srad.begin());
        range[(size_t)k] = make_int2(lo, std::max(lo, hi));
    }
    return 0;
}

static void str_density(tp_ctx *ctx, const StrAtom *d_atoms, const int2 *d_range, int LMAX, int K, double rmax,
                        double2 *surf)
{
    const int64_t n = (int64_t)K * 4 * LMAX * LMAX;
    hipLaunchKernelGGL(k_density_surface, dim3((unsigned)((n + kStrThreads - 1) / kStrThreads)), dim3(kStrThreads), 0,
                       ctx->stream, d_atoms, d_range, LMAX, K, rmax / K, surf);
}

int tp_density_surface(tp_ctx *ctx, const double *atoms, int natoms, int LMAX, int K, double rmax, double *surf)
{
    if (int rc = check_ctx(ctx)) return rc;
    if ((!atoms && natoms > 0) || natoms < 0 || !surf || LMAX < 1 || K < 1 || !(rmax > 0.0))
        return tp_fail(TP_E_ARG, "bad argument");
    std::vector<StrAtom> h;
    std::vector<int2> hr;
    if (int rc = str_atoms(atoms, natoms, K, rmax, h, hr)) return rc;
    const size_t n = (size_t)K * 4 * LMAX * LMAX;
    DevBuf<StrAtom> da;
    DevBuf<int2> dr;
    DevBuf<double2> f;
    if (natoms > 0)
        if (int rc = dev_alloc(&da.p, h.size())) return rc;
    if (int rc = dev_alloc(&dr.p, hr.size())) return rc;
    if (int rc = dev_alloc(&f.p, n)) return rc;
    if (natoms > 0)
        TP_HIP(hipMemcpyAsync(da.p, h.data(), sizeof(StrAtom) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipMemcpyAsync(dr.p, hr.data(), sizeof(int2) * hr.size(), hipMemcpyHostToDevice, ctx->stream));
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
    str_density(ctx, da.p, dr.p, LMAX, K, rmax, f.p);
    TP_HIP(hipGetLastError());
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
    TP_HIP(hipMemcpyAsync(surf, f.p, sizeof(double2) * n, hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->timing) {
        float tm = 0;
        if (hipEventElapsedTime(&tm, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->last_ms[0] = tm;
        for (int i = 1; i < 4; ++i) ctx->last_ms[i] = 0.0;
    }
    return 0;
}

int tp_absolute_square(tp_ctx *ctx, int LMAX, int K, const double *coeff, int64_t ld, double *out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!coeff || !out || ld < (int64_t)LMAX * LMAX) return tp_fail(TP_E_ARG, "bad argument");
    const PhTables *tcache = nullptr;
    if (int rc = ph_get(ctx, LMAX, K, false, &tcache)) return rc;
    const PhTables &t = *tcache;
    const size_t nq = (size_t)K * LMAX * LMAX, n = (size_t)K * 4 * LMAX * LMAX;
    DevBuf<double2> c, f;
    if (int rc = dev_alloc(&c.p, nq)) return rc;
    if (int rc = dev_alloc(&f.p, n)) return rc;
    const std::vector<double2> h = ph_pack(coeff, ld, K, LMAX);
    TP_HIP(hipMemcpyAsync(c.p, h.data(), sizeof(double2) * nq, hipMemcpyHostToDevice, ctx->stream));
    ph_synth(ctx, t, c.p, f.p);
    hipLaunchKernelGGL(k_abs_square, dim3((unsigned)((n + kStrThreads - 1) / kStrThreads)), dim3(kStrThreads), 0,
                       ctx->stream, f.p, (int64_t)n);
    ph_analysis(ctx, t, f.p, c.p);
    TP_HIP(hipGetLastError());
    std::vector<double2> r(nq);
    TP_HIP(hipMemcpyAsync(r.data(), c.p, sizeof(double2) * nq, hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    ph_unpack(r, K, LMAX, out, ld);
    return 0;
}

int tp_structure_from_atoms(tp_ctx *ctx, const double *atoms, int natoms, int LMAX, int K, double rmax,
                            double *density, double *fourier, double *intensity, int64_t ld)
{
    if (int rc = check_ctx(ctx)) return rc;
    if ((!atoms && natoms > 0) || natoms < 0 || !(rmax > 0.0) || K < 2 || LMAX < 1 || ld < (int64_t)LMAX * LMAX)
        return tp_fail(TP_E_ARG, "bad argument (natoms >= 0, rmax > 0, K >= 2, LMAX >= 1, ld >= LMAX^2)");
    std::vector<StrAtom> h;
    std::vector<int2> hr;
    if (int rc = str_atoms(atoms, natoms, K, rmax, h, hr)) return rc;
    const PhTables *tcache = nullptr;
    if (int rc = ph_get(ctx, LMAX, K, true, &tcache)) return rc;
    const PhTables &t = *tcache;
    const size_t nq = (size_t)K * LMAX * LMAX, n = (size_t)K * 4 * LMAX * LMAX;
    DevBuf<StrAtom> da;
    DevBuf<int2> dr;
    DevBuf<double2> f, cd, cf, ci;
    if (natoms > 0)
        if (int rc = dev_alloc(&da.p, h.size())) return rc;
    if (int rc = dev_alloc(&dr.p, hr.size())) return rc;
    if (int rc = dev_alloc(&f.p, n)) return rc;
    if (int rc = dev_alloc(&cd.p, nq)) return rc;
    if (int rc = dev_alloc(&cf.p, nq)) return rc;
    if (int rc = dev_alloc(&ci.p, nq)) return rc;
    hipStream_t st = ctx->stream;
    if (natoms > 0) TP_HIP(hipMemcpyAsync(da.p, h.data(), sizeof(StrAtom) * h.size(), hipMemcpyHostToDevice, st));
    TP_HIP(hipMemcpyAsync(dr.p, hr.data(), sizeof(int2) * hr.size(), hipMemcpyHostToDevice, st));
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[0], st));
    str_density(ctx, da.p, dr.p, LMAX, K, rmax, f.p);   // getDensity: sample
    ph_analysis(ctx, t, f.p, cd.p);                       //             and transform
    ph_bessel(ctx, t, cd.p, 1, cf.p);                     // forward
    ph_synth(ctx, t, cf.p, f.p);                          // absoluteSquare
    hipLaunchKernelGGL(k_abs_square, dim3((unsigned)((n + kStrThreads - 1) / kStrThreads)), dim3(kStrThreads), 0, st,
                       f.p, (int64_t)n);
    ph_analysis(ctx, t, f.p, ci.p);
    TP_HIP(hipGetLastError());
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[1], st));
    double *outs[3] = {density, fourier, intensity};
    const double2 *devs[3] = {cd.p, cf.p, ci.p};
    std::vector<double2> r[3];
    for (int o = 0; o < 3; ++o)
        if (outs[o]) {
            r[o].resize(nq);
            TP_HIP(hipMemcpyAsync(r[o].data(), devs[o], sizeof(double2) * nq, hipMemcpyDeviceToHost, st));
        }
    TP_HIP(hipStreamSynchronize(st));
    for (int o = 0; o < 3; ++o)
        if (outs[o]) ph_unpack(r[o], K, LMAX, outs[o], ld);
    if (ctx->timing) {
        float tm = 0;
        if (hipEventElapsedTime(&tm, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->last_ms[0] = tm;
        for (int i = 1; i < 4; ++i) ctx->last_ms[i] = 0.0;
    }
    return 0;
}

int tp_sraar(tp_ctx *ctx, int LMAX, int K, const double *intensity, int64_t ld, int iterations, double beta0,
             double beta_max, double tau, double cutoff_factor, const double *phases, uint64_t seed,
             double *coeff_out, double *diff_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!intensity || !coeff_out || ld < (int64_t)LMAX * LMAX || iterations < 0)
        return tp_fail(TP_E_ARG, "bad argument");
    const PhTables *tcache = nullptr;
    if (int rc = ph_get(ctx, LMAX, K, true, &tcache)) return rc;
    const PhTables &t = *tcache;
    hipStream_t st = ctx->stream;
    const int NP = 4 * LMAX * LMAX;
    const int64_t n = (int64_t)K * NP;
    const int nq = LMAX * LMAX;
    const int Koff = (int)std::ceil(cutoff_factor * K);
    const unsigned nb = (unsigned)((n + kPhThreads - 1) / kPhThreads);
    DevBuf<double2> C1, C2, F, dens, tmpu, tmp1, tmp2;
    DevBuf<double> amp, ph, part, diff;
    int rc = 0;
    if ((rc = dev_alloc(&C1.p, (size_t)K * nq)) || (rc = dev_alloc(&C2.p, (size_t)K * nq)) ||
        (rc = dev_alloc(&F.p, (size_t)n)) || (rc = dev_alloc(&dens.p, (size_t)n)) || (rc = dev_alloc(&tmpu.p, (size_t)n)) ||
        (rc = dev_alloc(&tmp1.p, (size_t)n)) || (rc = dev_alloc(&tmp2.p, (size_t)n)) || (rc = dev_alloc(&amp.p, (size_t)n)) ||
        (rc = dev_alloc(&ph.p, (size_t)n)) || (rc = dev_alloc(&part.p, (size_t)2 * nb)) ||
        (rc = dev_alloc(&diff.p, (size_t)std::max(iterations, 1))))
        return rc;
    {
        const std::vector<double2> h = ph_pack(intensity, ld, K, LMAX);
        TP_HIP(hipMemcpyAsync(C1.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, st));
        std::vector<double> hp((size_t)n);
        if (phases) {
            std::memcpy(hp.data(), phases, sizeof(double) * (size_t)n);
        } else {   // pi * 2 * rand per sample (phases.jl:48-52)
            std::mt19937_64 g(seed);
            std::uniform_real_distribution<double> u(0.0, 1.0);
            for (auto &v : hp) v = tph::kPi * 2 * u(g);
        }
        TP_HIP(hipMemcpyAsync(ph.p, hp.data(), sizeof(double) * hp.size(), hipMemcpyHostToDevice, st));
        TP_HIP(hipStreamSynchronize(st));   // the host staging vectors end here
    }
    // B(x) -> out: forward (analysis, Bessel, synthesis), modulus, backward
    auto projB = [&](const double2 *x, double2 *out) {
        ph_analysis(ctx, t, x, C1.p);
        ph_bessel(ctx, t, C1.p, 1, C2.p);
        ph_synth(ctx, t, C2.p, F.p);
        hipLaunchKernelGGL(k_raar_modulus, dim3(nb), dim3(kPhThreads), 0, st, F.p, amp.p, n);
        ph_analysis(ctx, t, F.p, C1.p);
        ph_bessel(ctx, t, C1.p, 0, C2.p);
        ph_synth(ctx, t, C2.p, out);
    };
    // amplitudes and random phases; density = backward(fourier); tmp1 = 2 B(density) - density
    ph_synth(ctx, t, C1.p, tmp2.p);
    hipLaunchKernelGGL(k_raar_init, dim3(nb), dim3(kPhThreads), 0, st, tmp2.p, ph.p, n, amp.p, F.p);
    ph_analysis(ctx, t, F.p, C1.p);
    ph_bessel(ctx, t, C1.p, 0, C2.p);
    ph_synth(ctx, t, C2.p, dens.p);
    projB(dens.p, tmp2.p);
    hipLaunchKernelGGL(k_raar_reflect, dim3(nb), dim3(kPhThreads), 0, st, tmp2.p, dens.p, n, tmp1.p);
    TP_HIP(hipGetLastError());
    double2 *d = dens.p, *u = tmpu.p;
    for (int it = 1; it <= iterations; ++it) {
        const double e3 = std::exp(std::pow(-it / tau, 3.0));
        const double beta = e3 * beta0 + (1.0 - e3) * beta_max;
        hipLaunchKernelGGL(k_raar_update, dim3(nb), dim3(kPhThreads), 0, st, tmp1.p, d, n, NP, Koff, beta, u, part.p);
        hipLaunchKernelGGL(k_raar_diff, dim3(1), dim3(64), 0, st, part.p, (int)nb, diff.p, it - 1);
        projB(u, tmp2.p);
        std::swap(d, u);   // density = tmp_u
        hipLaunchKernelGGL(k_raar_reflect, dim3(nb), dim3(kPhThreads), 0, st, tmp2.p, d, n, tmp1.p);
        TP_HIP(hipGetLastError());
    }
    // density = A(tmp2) (tmp2 = tmp1 when no iteration ran, phases.jl:59); its coefficients
    hipLaunchKernelGGL(k_raar_support, dim3(nb), dim3(kPhThreads), 0, st, iterations > 0 ? tmp2.p : tmp1.p, n, NP, Koff,
                       u);
    ph_analysis(ctx, t, u, C1.p);
    TP_HIP(hipGetLastError());
    std::vector<double2> h((size_t)K * nq);
    TP_HIP(hipMemcpyAsync(h.data(), C1.p, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, st));
    if (diff_out && iterations > 0)
        TP_HIP(hipMemcpyAsync(diff_out, diff.p, sizeof(double) * iterations, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    ph_unpack(h, K, LMAX, coeff_out, ld);
    return 0;
}

// ---------------------------------------------------------------------------
// synthetic pictures (pointsPerOrientation, datagen.jl:88-144) and the
// histogram generation driver over them (generateHistogram, datagen.jl:363-415)
// ---------------------------------------------------------------------------
// The shells' surfaces (getSurfaceVolume) on the device and the envelope bound
// maxIntensity = 1.3 max Re surf (datagen.jl:89-92 for a SurfaceVolume).
static int pic_surfaces(tp_ctx *ctx, int LMAX, int K, const double *coeff, int64_t ld, DevBuf<double2> &surf,
                        double &maxI)
{
    const PhTables *tcache = nullptr;
    if (int rc = ph_get(ctx, LMAX, K, false, &tcache)) return rc;
    const PhTables &t = *tcache;
    const size_t NP = (size_t)4 * LMAX * LMAX;
    DevBuf<double2> c;
    if (int rc = dev_alloc(&c.p, (size_t)K * LMAX * LMAX)) return rc;
    if (int rc = dev_alloc(&surf.p, (size_t)K * NP)) return rc;
    const std::vector<double2> h = ph_pack(coeff, ld, K, LMAX);
    TP_HIP(hipMemcpyAsync(c.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    ph_synth(ctx, t, c.p, surf.p);
    TP_HIP(hipGetLastError());
    std::vector<double2> hs((size_t)K * NP);
    TP_HIP(hipMemcpyAsync(hs.data(), surf.p, sizeof(double2) * hs.size(), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    double mx = -INFINITY;
    for (const double2 &v : hs) mx = std::max(mx, v.x);
    maxI = mx * 1.3;
    return 0;
}

// Surfaces of the noise volume of get_noise_volume (datagen.jl:296-304): every
// coefficient deleted, c00(k) = gaussian_distribution(k dr; 0, sigma)
// (utilities.jl:129-133).
static int pic_noise_surfaces(tp_ctx *ctx, int LMAX, int K, double rmax, double sigma, DevBuf<double2> &surf,
                              double &maxI)
{
    std::vector<double> c((size_t)2 * K * LMAX * LMAX, 0.0);
    const double s2 = 2.0 * sigma * sigma;
    for (int k = 1; k <= K; ++k) {
        const double q = k * (rmax / K);
        c[(size_t)2 * (k - 1) * LMAX * LMAX] = 1 / std::sqrt(tph::kPi * s2) * std::exp(-(q * q) / s2);
    }
    return pic_surfaces(ctx, LMAX, K, c.data(), (int64_t)LMAX * LMAX, surf, maxI);
}

static int pic_check(const tp_picture_params *p, int LMAX, int K, double rmax, int npictures)
{
    if (!p) return tp_fail(TP_E_ARG, "null parameters");
    if (p->particles < 1 || p->noise_photons < 0 || (p->noise_photons > 0 && !(p->noise_sigma > 0.0)))
        return tp_fail(TP_E_ARG, "need particles >= 1 and, with noise photons, noise_sigma > 0");
    if (LMAX < 1 || LMAX > kPhMaxLMAX || K < 1 || !(rmax > 0.0)) return tp_fail(TP_E_ARG, "need 1 <= LMAX <= 48, K >= 1, rmax > 0");
    if (npictures < 0 || p->photons_per_image < 0 || p->capacity < 1 || p->incident_photon_variance < 0)
        return tp_fail(TP_E_ARG, "bad picture counts");
    // photon indices 3 i + 1 of a segment stay below 2^31 (a count normal beyond 20 deviations never comes)
    if ((double)p->photons_per_image + 20.0 * p->incident_photon_variance > (double)(1 << 28) ||
        p->noise_photons > (1 << 28))
        return tp_fail(TP_E_UNSUPPORTED, "more than 2^28 incident photons per segment");
    if (!(p->envelope_sigma > 0.0) || !(p->qcut > 0.0)) return tp_fail(TP_E_ARG, "need qcut, envelope_sigma > 0");
    return 0;
}

static PicArgs pic_args(const tp_picture_params *p, int LMAX, int K, double rmax, const double2 *surf, double maxI,
                        const double2 *nsurf, double nmaxI, int first)
{
    PicArgs a;
    a.surf = surf;
    a.nsurf = nsurf;
    a.LMAX = LMAX;
    a.K = K;
    a.dr = rmax / K;
    a.maxI = maxI;
    a.nmaxI = nmaxI;
    a.qcut = p->qcut;
    a.sigma = p->envelope_sigma;
    a.nsigma = 1.05 * p->noise_sigma;   // datagen.jl:404
    a.lambda = p->lambda;
    a.bw = p->beamstop_width;
    a.n_inc = p->photons_per_image;
    a.n_noise = p->beamstop_only ? 0 : p->noise_photons;
    a.variance = p->incident_photon_variance;
    a.particles = p->particles;
    a.beamstop_only = p->beamstop_only ? 1 : 0;
    a.cap = p->capacity;
    a.first = first;
    a.seed = p->seed;
    return a;
}

int tp_pictures(tp_ctx *ctx, int LMAX, int K, double rmax, const double *coeff, int64_t ld, const tp_picture_params *p,
                int npictures, int first_picture, double *photons, int32_t *counts, double *rotations)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (int rc = pic_check(p, LMAX, K, rmax, npictures)) return rc;
    if (!coeff || !photons || !counts || ld < (int64_t)LMAX * LMAX) return tp_fail(TP_E_ARG, "bad argument");
    if (npictures == 0) return 0;
    DevBuf<double2> surf, nsurf;
    double maxI = 0.0, nmaxI = 0.0;
    if (int rc = pic_surfaces(ctx, LMAX, K, coeff, ld, surf, maxI)) return rc;
    if (p->noise_photons > 0 && !p->beamstop_only)
        if (int rc = pic_noise_surfaces(ctx, LMAX, K, rmax, p->noise_sigma, nsurf, nmaxI)) return rc;
    DevBuf<double> dph, drot;
    DevBuf<int> dcnt;
    if (int rc = dev_alloc(&dph.p, (size_t)npictures * p->capacity * 3)) return rc;
    if (int rc = dev_alloc(&dcnt.p, (size_t)npictures)) return rc;
    if (rotations)
        if (int rc = dev_alloc(&drot.p, (size_t)npictures * 9)) return rc;
    hipStream_t st = ctx->stream;
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[0], st));
    hipLaunchKernelGGL(k_pictures, dim3((unsigned)npictures), dim3(kPicThreads), 0, st,
                       pic_args(p, LMAX, K, rmax, surf.p, maxI, nsurf.p, nmaxI, first_picture), dph.p, dcnt.p, drot.p);
    TP_HIP(hipGetLastError());
    if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[1], st));
    TP_HIP(hipMemcpyAsync(photons, dph.p, sizeof(double) * (size_t)npictures * p->capacity * 3, hipMemcpyDeviceToHost, st));
    TP_HIP(hipMemcpyAsync(counts, dcnt.p, sizeof(int) * (size_t)npictures, hipMemcpyDeviceToHost, st));
    if (rotations) TP_HIP(hipMemcpyAsync(rotations, drot.p, sizeof(double) * 9 * (size_t)npictures, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    if (ctx->timing) {
        float tm = 0;
        if (hipEventElapsedTime(&tm, ctx->ev[0], ctx->ev[1]) == hipSuccess) ctx->last_ms[0] = tm;
        for (int i = 1; i < 4; ++i) ctx->last_ms[i] = 0.0;
    }
    return 0;
}

int tp_hist_add_generated(tp_hist *h, int LMAX, int K, double rmax, const double *coeff, int64_t ld,
                          const tp_picture_params *p, int npictures, int first_picture, int64_t *overflow)
{
    if (!h) return tp_fail(TP_E_ARG, "null histogram");
    tp_ctx *ctx = h->ctx;
    if (int rc = check_ctx(ctx)) return rc;
    if (int rc = pic_check(p, LMAX, K, rmax, npictures)) return rc;
    if (!coeff || ld < (int64_t)LMAX * LMAX) return tp_fail(TP_E_ARG, "bad argument");
    if (p->capacity > kHistMaxPhotons) return tp_fail(TP_E_UNSUPPORTED, "capacity above 512 photons per picture");
    if (npictures == 0) return 0;
    DevBuf<double2> surf, nsurf;
    double maxI = 0.0, nmaxI = 0.0;
    if (int rc = pic_surfaces(ctx, LMAX, K, coeff, ld, surf, maxI)) return rc;
    if (p->noise_photons > 0 && !p->beamstop_only)
        if (int rc = pic_noise_surfaces(ctx, LMAX, K, rmax, p->noise_sigma, nsurf, nmaxI)) return rc;
    const int batch = std::min(npictures, 8192);
    DevBuf<double> dph;
    DevBuf<int> dcnt;
    if (int rc = dev_alloc(&dph.p, (size_t)batch * p->capacity * 3)) return rc;
    if (int rc = dev_alloc(&dcnt.p, (size_t)batch)) return rc;
    hipStream_t st = ctx->stream;
    std::vector<int> hc(batch);
    int64_t over = 0;
    float t_gen = 0.f, t_hist = 0.f;
    for (int b0 = 0; b0 < npictures; b0 += batch) {
        const int nb = std::min(batch, npictures - b0);
        if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[0], st));
        hipLaunchKernelGGL(k_pictures, dim3((unsigned)nb), dim3(kPicThreads), 0, st,
                           pic_args(p, LMAX, K, rmax, surf.p, maxI, nsurf.p, nmaxI, first_picture + b0), dph.p, dcnt.p,
                           (double *)nullptr);
        TP_HIP(hipGetLastError());
        if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[1], st));
        // a picture that accepted more than capacity photons is skipped whole
        hipLaunchKernelGGL(k_hist_pictures, dim3((unsigned)nb), dim3(kHistThreads), 0, st, dph.p, (const int64_t *)nullptr,
                           dcnt.p, p->capacity, h->dq, h->N, h->K2, h->K3, h->d_n1, h->d_n2, h->d_n3);
        TP_HIP(hipGetLastError());
        if (ctx->timing) TP_HIP(hipEventRecord(ctx->ev[2], st));
        TP_HIP(hipMemcpyAsync(hc.data(), dcnt.p, sizeof(int) * (size_t)nb, hipMemcpyDeviceToHost, st));
        TP_HIP(hipStreamSynchronize(st));
        for (int i = 0; i < nb; ++i) over += hc[i] > p->capacity ? 1 : 0;
        if (ctx->timing) {
            float a = 0.f, c = 0.f;
            if (hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]) == hipSuccess) t_gen += a;
            if (hipEventElapsedTime(&c, ctx->ev[1], ctx->ev[2]) == hipSuccess) t_hist += c;
        }
    }
    if (ctx->timing) {
        ctx->last_ms[0] = t_gen;
        ctx->last_ms[1] = t_hist;
        ctx->last_ms[2] = ctx->last_ms[3] = 0.0;
    }
    h->pictures += npictures - over;
    if (overflow) *overflow = over;
    return 0;
}

void tp_state_destroy(tp_state *s)
{
    if (!s) return;
    if (s->b && s->b->ctx) (void)hipSetDevice(s->b->ctx->device);
    void *ptrs[] = {s->d_cur, s->d_prop, s->d_scal, s->d_step, s->d_sel};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
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
        (rc = dev_alloc(&s->d_sel, 4))) {
        tp_state_destroy(s);
        return rc;
    }
    if ((rc = upload_coeff(ctx, b, coeff, ld, nshell, s->d_cur, nullptr))) {
        tp_state_destroy(s);
        return rc;
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        tp_state_destroy(s);
        return tp_fail(TP_E_HIP, "state upload failed");
    }
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
    McArgs mc = mc_none();
    mc.seqp = reinterpret_cast<uint64_t *>(ctx->d_res + 1);
    mc.seq = ++ctx->res_seq;
    if (int rc = enqueue_eval(ctx, b, r, s->d_cur, nullptr, nullptr, 0, nullptr, ctx->d_res, mc)) return rc;
    if (int rc = wait_result(ctx, mc.seq)) return rc;
    *E_out = ctx->h_res[0];
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
    if (int rc = enqueue_eval(ctx, b, nullptr, s->d_cur, nullptr, nullptr, 0, b->d_c3, nullptr, mc_none())) return rc;
    TP_HIP(hipMemcpyAsync(c3_out, b->d_c3, sizeof(double) * (size_t)b->klength * b->R, hipMemcpyDeviceToHost,
                          ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

static int anneal_check(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, const tp_state *s, const tp_anneal_params *p)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!b || !r || !s || !p) return tp_fail(TP_E_ARG, "null argument");
    if (b->ctx != ctx || r->b != b || s->b != b) return tp_fail(TP_E_STATE, "objects belong to another context/basis");
    if (p->nsteps < 0 || p->ntemp < 0 || !(p->stepsizefactor > 0) || !(p->temperature_decay > 0))
        return tp_fail(TP_E_ARG, "invalid annealing parameters");
    if (b->L < 2) return tp_fail(TP_E_ARG, "annealing needs L >= 2");
    if (b->L > 62) return tp_fail(TP_E_UNSUPPORTED, "annealing moves need 2L+1 <= 64");
    return 0;
}

static McArgs anneal_mc(const tp_basis *b, tp_state *s, const tp_anneal_params *p)
{
    McArgs mc = mc_none();
    mc.seed = p->seed;
    mc.decay = p->temperature_decay;
    mc.factor = p->stepsizefactor;
    mc.nl = b->L / 2;
    mc.scal = s->d_scal;
    mc.stepsizes = s->d_step;
    mc.sel = s->d_sel;
    return mc;
}

static void launch_propose(tp_ctx *ctx, const tp_basis *b, tp_state *s, uint64_t seed, uint32_t step, uint32_t tag,
                           const int *stopped)
{
    hipLaunchKernelGGL(k_propose, dim3((unsigned)(b->L / 2), (unsigned)((s->nshell + 3) / 4)), dim3(256), 0,
                       ctx->stream, s->d_cur, s->d_prop, s->d_sel, s->nshell, b->n_c, b->d_coff, s->d_step, seed, step,
                       tag, stopped);
}

int tp_anneal_begin(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, const tp_anneal_params *p,
                    tp_anneal_status *st)
{
    if (int rc = anneal_check(ctx, b, r, s, p)) return rc;
    if (!st) return tp_fail(TP_E_ARG, "null status");
    const int nl = b->L / 2;
    hipStream_t stream = ctx->stream;
    std::memset(st, 0, sizeof(*st));
    st->nl = nl;
    for (int l = 2; l <= b->L; l += 2) st->stepsizes[l / 2 - 1] = p->initial_stepsize * (l - 1);   // determination.jl:295
    TP_HIP(hipMemcpyAsync(s->d_step, st->stepsizes, sizeof(double) * nl, hipMemcpyHostToDevice, stream));
    TP_HIP(hipMemsetAsync(s->d_sel, 0, sizeof(int) * 4, stream));
    double *d_temp = nullptr;
    if (int rc = dev_alloc(&d_temp, (size_t)std::max(1, p->ntemp))) return rc;
    McArgs mc = anneal_mc(b, s, p);
    mc.temp = d_temp;
    // initial energy (determination.jl:298) -> scal[0]
    mc.mode = 2;
    int rc = enqueue_eval(ctx, b, r, s->d_cur, s->d_prop, s->d_sel, 0, nullptr, b->d_out, mc);
    // temperature from `ntemp` proposals of the start structure (calculate_temperature, :244-261)
    double T0 = 0.0;
    if (!rc && p->initial_temperature_factor > 0.0 && p->ntemp > 1) {
        mc.mode = 1;
        for (int i = 0; i < p->ntemp && !rc; ++i) {
            launch_propose(ctx, b, s, p->seed, (uint32_t)i, TP_TAG_TEMP, nullptr);
            mc.step = (uint32_t)i;
            rc = enqueue_eval(ctx, b, r, s->d_cur, s->d_prop, s->d_sel, 1, nullptr, b->d_out, mc);
        }
        std::vector<double> et(p->ntemp);
        if (!rc && (hipMemcpyAsync(et.data(), d_temp, sizeof(double) * p->ntemp, hipMemcpyDeviceToHost, stream) !=
                        hipSuccess ||
                    hipStreamSynchronize(stream) != hipSuccess))
            rc = tp_fail(TP_E_HIP, "temperature probe read failed");
        if (!rc) {
            double m = 0;
            for (double v : et) m += v;
            m /= p->ntemp;
            double v2 = 0;
            for (double v : et) v2 += (v - m) * (v - m);
            T0 = std::sqrt(v2 / (p->ntemp - 1)) * p->initial_temperature_factor;
        }
    }
    (void)hipFree(d_temp);
    if (rc) return rc;
    TP_HIP(hipMemcpyAsync(s->d_scal + 1, &T0, sizeof(double), hipMemcpyHostToDevice, stream));
    double E0 = 0;
    TP_HIP(hipMemcpyAsync(&E0, s->d_scal, sizeof(double), hipMemcpyDeviceToHost, stream));
    TP_HIP(hipStreamSynchronize(stream));
    st->E = st->E0 = E0;
    st->T = st->T0 = T0;
    return 0;
}

int tp_anneal_steps(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, const tp_anneal_params *p,
                    tp_anneal_status *st, int max_steps, double min_stepsize, double *E_trace, double *Enew_trace,
                    int32_t *acc_trace, int32_t *steps_done)
{
    if (steps_done) *steps_done = 0;
    if (int rc = anneal_check(ctx, b, r, s, p)) return rc;
    if (!st) return tp_fail(TP_E_ARG, "null status");
    const int nl = b->L / 2;
    if (st->nl != nl) return tp_fail(TP_E_STATE, "status was made for another L");
    if (max_steps < 0 || st->step < 0 || st->step + (int64_t)max_steps > (int64_t)INT32_MAX)
        return tp_fail(TP_E_ARG, "step count out of range");
    // the loop condition is checked before each step (determination.jl:322)
    if (min_stepsize > 0.0 && st->stepsizes[nl - 1] <= min_stepsize) st->stopped = 1;
    if (st->stopped || max_steps == 0) return 0;
    hipStream_t stream = ctx->stream;
    const int32_t ctl0[4] = {0, 0, (int32_t)st->step, st->naccepted};   // sel, stopped, last step, accepted
    const double sc0[2] = {st->E, st->T};
    TP_HIP(hipMemcpyAsync(s->d_sel, ctl0, sizeof(ctl0), hipMemcpyHostToDevice, stream));
    TP_HIP(hipMemcpyAsync(s->d_scal, sc0, sizeof(sc0), hipMemcpyHostToDevice, stream));
    TP_HIP(hipMemcpyAsync(s->d_step, st->stepsizes, sizeof(double) * nl, hipMemcpyHostToDevice, stream));
    double *d_tr = nullptr;
    int *d_acc = nullptr;
    int rc = 0;
    if ((rc = dev_alloc(&d_tr, 2 * (size_t)max_steps)) || (rc = dev_alloc(&d_acc, (size_t)max_steps))) {
        if (d_tr) (void)hipFree(d_tr);
        return rc;
    }
    McArgs mc = anneal_mc(b, s, p);
    mc.mode = 3;
    mc.trE = d_tr;
    mc.trEn = d_tr + max_steps;
    mc.trA = d_acc;
    mc.tr0 = (uint32_t)st->step;
    mc.minstep = min_stepsize;
    mc.ctl = s->d_sel + 1;
    // the hot loop: propose, evaluate the proposal buffer, Metropolis in k_final;
    // with a stopping rule the host looks at the flag every kCheck steps
    constexpr int kCheck = 256;
    const uint32_t first = (uint32_t)st->step + 1, last = (uint32_t)(st->step + max_steps);
    for (uint32_t i = first; i <= last && !rc; ++i) {
        launch_propose(ctx, b, s, p->seed, i, TP_TAG_MOVE, min_stepsize > 0.0 ? mc.ctl : nullptr);
        mc.step = i;
        rc = enqueue_eval(ctx, b, r, s->d_cur, s->d_prop, s->d_sel, 1, nullptr, b->d_out, mc);
        if (!rc && min_stepsize > 0.0 && (i - first + 1) % kCheck == 0 && i < last) {
            int32_t stopped = 0;
            if (hipMemcpyAsync(&stopped, mc.ctl, sizeof(int32_t), hipMemcpyDeviceToHost, stream) != hipSuccess ||
                hipStreamSynchronize(stream) != hipSuccess)
                rc = tp_fail(TP_E_HIP, "stopping-rule flag read failed");
            if (stopped) break;
        }
    }
    if (!rc && hipGetLastError() != hipSuccess) rc = tp_fail(TP_E_HIP, "annealing launch failed");
    int32_t ctl[4] = {0, 0, 0, 0};
    double sc[2] = {0, 0};
    if (!rc && (hipMemcpyAsync(ctl, s->d_sel, sizeof(ctl), hipMemcpyDeviceToHost, stream) != hipSuccess ||
                hipMemcpyAsync(sc, s->d_scal, sizeof(sc), hipMemcpyDeviceToHost, stream) != hipSuccess ||
                hipMemcpyAsync(st->stepsizes, s->d_step, sizeof(double) * nl, hipMemcpyDeviceToHost, stream) !=
                    hipSuccess ||
                hipStreamSynchronize(stream) != hipSuccess))
        rc = tp_fail(TP_E_HIP, "annealing state read failed");
    const int32_t taken = rc ? 0 : (int32_t)(ctl[2] - st->step);
    if (!rc && taken > 0) {
        if ((E_trace && hipMemcpy(E_trace, d_tr, sizeof(double) * taken, hipMemcpyDeviceToHost) != hipSuccess) ||
            (Enew_trace &&
             hipMemcpy(Enew_trace, d_tr + max_steps, sizeof(double) * taken, hipMemcpyDeviceToHost) != hipSuccess) ||
            (acc_trace && hipMemcpy(acc_trace, d_acc, sizeof(int32_t) * taken, hipMemcpyDeviceToHost) != hipSuccess))
            rc = tp_fail(TP_E_HIP, "trace read failed");
    }
    (void)hipFree(d_tr);
    (void)hipFree(d_acc);
    if (rc) return rc;
    if (ctl[0]) std::swap(s->d_cur, s->d_prop);   // the accepted structure is current again
    st->step = ctl[2];
    st->naccepted = ctl[3];
    st->stopped = ctl[1];
    st->E = sc[0];
    st->T = sc[1];
    if (min_stepsize > 0.0 && st->stepsizes[nl - 1] <= min_stepsize) st->stopped = 1;
    if (steps_done) *steps_done = taken;
    return 0;
}

int tp_anneal(tp_ctx *ctx, const tp_basis *b, const tp_ref *r, tp_state *s, const tp_anneal_params *p,
              double *E_trace, double *Enew_trace, int32_t *acc_trace, double *scalars)
{
    tp_anneal_status st;
    if (int rc = tp_anneal_begin(ctx, b, r, s, p, &st)) return rc;
    if (int rc = tp_anneal_steps(ctx, b, r, s, p, &st, p->nsteps, 0.0, E_trace, Enew_trace, acc_trace, nullptr))
        return rc;
    if (scalars) {
        scalars[0] = st.E0; scalars[1] = st.T0; scalars[2] = st.E; scalars[3] = st.T;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Checkpoints (saveState / rotation_search's resume, determination.jl:154-278)
// File: "TPCKPT\0\0", u32 version, u32 header bytes, the header fields below
// in this order (little-endian, as written by this x86-64 host code), the
// coefficients (nshell * ld complex128), then u64 FNV-1a of all bytes before.
// ---------------------------------------------------------------------------
}   // extern "C": the serialisation helpers below are C++
static const char kCkptMagic[8] = {'T', 'P', 'C', 'K', 'P', 'T', 0, 0};
constexpr uint32_t kCkptVersion = 1;

struct CkptWriter {
    std::vector<unsigned char> buf;
    template <class T>
    void put(const T &v)
    {
        const unsigned char *p = reinterpret_cast<const unsigned char *>(&v);
        buf.insert(buf.end(), p, p + sizeof(T));
    }
};

struct CkptReader {
    const unsigned char *p, *end;
    bool ok = true;
    template <class T>
    T get()
    {
        T v{};
        if ((size_t)(end - p) < sizeof(T)) {
            ok = false;
            return v;
        }
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
};

static uint64_t fnv1a(const unsigned char *p, size_t n)
{
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

static void ckpt_header(CkptWriter &w, const tp_checkpoint_info *in)
{
    const tp_anneal_params &pa = in->params;
    const tp_anneal_status &s = in->status;
    w.put(pa.nsteps); w.put(pa.ntemp); w.put(pa.initial_stepsize); w.put(pa.stepsizefactor);
    w.put(pa.initial_temperature_factor); w.put(pa.temperature_decay); w.put(pa.seed);
    w.put(s.step); w.put(s.E); w.put(s.T); w.put(s.E0); w.put(s.T0);
    for (int i = 0; i < TP_MAX_LBLOCKS; ++i) w.put(s.stepsizes[i]);
    w.put(s.nl); w.put(s.naccepted); w.put(s.stopped);
    w.put(in->min_stepsize); w.put(in->lambda); w.put(in->dq);
    w.put(in->L); w.put(in->LMAX); w.put(in->N); w.put(in->K); w.put(in->k3lo); w.put(in->k3hi);
    w.put(in->nshell); w.put(in->ld);
}

static void ckpt_read_header(CkptReader &r, tp_checkpoint_info *o)
{
    std::memset(o, 0, sizeof(*o));
    tp_anneal_params &pa = o->params;
    tp_anneal_status &s = o->status;
    pa.nsteps = r.get<int>(); pa.ntemp = r.get<int>(); pa.initial_stepsize = r.get<double>();
    pa.stepsizefactor = r.get<double>(); pa.initial_temperature_factor = r.get<double>();
    pa.temperature_decay = r.get<double>(); pa.seed = r.get<uint64_t>();
    s.step = r.get<int64_t>(); s.E = r.get<double>(); s.T = r.get<double>(); s.E0 = r.get<double>();
    s.T0 = r.get<double>();
    for (int i = 0; i < TP_MAX_LBLOCKS; ++i) s.stepsizes[i] = r.get<double>();
    s.nl = r.get<int32_t>(); s.naccepted = r.get<int32_t>(); s.stopped = r.get<int32_t>();
    o->min_stepsize = r.get<double>(); o->lambda = r.get<double>(); o->dq = r.get<double>();
    o->L = r.get<int32_t>(); o->LMAX = r.get<int32_t>(); o->N = r.get<int32_t>(); o->K = r.get<int32_t>();
    o->k3lo = r.get<int32_t>(); o->k3hi = r.get<int32_t>(); o->nshell = r.get<int32_t>(); o->ld = r.get<int32_t>();
}

extern "C" {
int tp_checkpoint_save(const char *path, const tp_checkpoint_info *info, const double *coeff)
{
    if (!path || !info || (!coeff && info->nshell > 0)) return tp_fail(TP_E_ARG, "null argument");
    if (info->nshell < 0 || info->ld < 0 || (int64_t)info->ld < (int64_t)info->LMAX * info->LMAX ||
        info->status.nl < 0 || info->status.nl > TP_MAX_LBLOCKS)
        return tp_fail(TP_E_ARG, "checkpoint shape out of range (nshell >= 0, ld >= LMAX^2, nl <= 32)");
    CkptWriter hw;
    ckpt_header(hw, info);
    CkptWriter w;
    w.buf.insert(w.buf.end(), kCkptMagic, kCkptMagic + 8);
    w.put(kCkptVersion);
    w.put((uint32_t)hw.buf.size());
    w.buf.insert(w.buf.end(), hw.buf.begin(), hw.buf.end());
    const size_t nd = 2 * (size_t)info->nshell * (size_t)info->ld;
    const unsigned char *cp = reinterpret_cast<const unsigned char *>(coeff);
    if (nd) w.buf.insert(w.buf.end(), cp, cp + nd * sizeof(double));
    w.put(fnv1a(w.buf.data(), w.buf.size()));
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return tp_fail(TP_E_ARG, "cannot open " + tmp + " for writing");
    const bool ok = std::fwrite(w.buf.data(), 1, w.buf.size(), f) == w.buf.size();
    if (std::fclose(f) != 0 || !ok) {
        std::remove(tmp.c_str());
        return tp_fail(TP_E_ARG, "short write to " + tmp);
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return tp_fail(TP_E_ARG, std::string("cannot rename the checkpoint to ") + path);
    }
    return 0;
}

int tp_checkpoint_load(const char *path, tp_checkpoint_info *info, double *coeff, int64_t cap_doubles)
{
    if (!path || !info) return tp_fail(TP_E_ARG, "null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return tp_fail(TP_E_ARG, std::string("cannot open checkpoint ") + path);
    std::vector<unsigned char> buf;
    unsigned char chunk[1 << 16];
    size_t n;
    while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    std::fclose(f);
    if (buf.size() < 8 + 4 + 4 + 8 || std::memcmp(buf.data(), kCkptMagic, 8) != 0)
        return tp_fail(TP_E_ARG, std::string(path) + " is not a checkpoint of this library");
    uint64_t sum;
    std::memcpy(&sum, buf.data() + buf.size() - 8, 8);
    if (fnv1a(buf.data(), buf.size() - 8) != sum) return tp_fail(TP_E_ARG, std::string(path) + ": checksum mismatch");
    CkptReader r{buf.data() + 8, buf.data() + buf.size() - 8};
    const uint32_t ver = r.get<uint32_t>(), hlen = r.get<uint32_t>();
    if (ver != kCkptVersion) return tp_fail(TP_E_ARG, std::string(path) + ": unknown checkpoint version");
    CkptReader h{r.p, r.p + std::min<size_t>(hlen, (size_t)(r.end - r.p))};
    ckpt_read_header(h, info);
    if (!h.ok || !r.ok) return tp_fail(TP_E_ARG, std::string(path) + ": truncated header");
    r.p += hlen;
    if (info->nshell < 0 || info->ld < 0) return tp_fail(TP_E_ARG, std::string(path) + ": bad shape");
    const size_t nd = 2 * (size_t)info->nshell * (size_t)info->ld;
    if ((size_t)(r.end - r.p) != nd * sizeof(double))
        return tp_fail(TP_E_ARG, std::string(path) + ": coefficient block has the wrong size");
    if (coeff) {
        if ((int64_t)nd > cap_doubles) return tp_fail(TP_E_ARG, "coefficient buffer too small for the checkpoint");
        std::memcpy(coeff, r.p, nd * sizeof(double));
    }
    return 0;
}

int tp_load_histograms(const double *c3_full, int NA, int NB, int Kf, int K3, double *c3_out)
{
    if (!c3_full || !c3_out) return tp_fail(TP_E_ARG, "null argument");
    if (NA < 1 || Kf < 1 || K3 < 1 || K3 > Kf || (NB != NA && NB != 2 * NA))
        return tp_fail(TP_E_ARG, "histogram must be N x N or N x 2N x Kf^3 with 1 <= K3 <= Kf");
    const int64_t N = NA, R = 2 * N * N, Rf = (int64_t)NA * NB;
    const int64_t nout = R * K3 * K3 * K3;
    // doubling of an old N x N histogram: only k1 >= k2 >= k3 is filled (datagen.jl:445-458),
    // every other slice stays 0 and is clamped below
    for (int k1 = 1; k1 <= K3; ++k1)
        for (int k2 = 1; k2 <= K3; ++k2)
            for (int k3 = 1; k3 <= K3; ++k3) {
                const double *src = c3_full + Rf * ((k3 - 1) + (int64_t)Kf * ((k2 - 1) + (int64_t)Kf * (k1 - 1)));
                double *dst = c3_out + R * ((k3 - 1) + (int64_t)K3 * ((k2 - 1) + (int64_t)K3 * (k1 - 1)));
                if (NB == 2 * NA) {
                    std::memcpy(dst, src, sizeof(double) * R);
                } else if (k1 >= k2 && k2 >= k3) {
                    std::memcpy(dst, src, sizeof(double) * Rf);
                    std::memcpy(dst + Rf, src, sizeof(double) * Rf);
                } else {
                    std::fill(dst, dst + R, 0.0);
                }
            }
    // clamp (:460-462); the cut keeps K3 shells, then the global normalisation (:465-468)
    double tot = 0.0;
    for (int64_t i = 0; i < nout; ++i) {
        if (std::isnan(c3_out[i])) return tp_fail(TP_E_ARG, "histogram holds NaN");   // max.(NaN, 1e-30) is NaN
        const double v = c3_out[i] > 1e-30 ? c3_out[i] : 1e-30;
        c3_out[i] = v;
        tot += std::fabs(v);
    }
    for (int64_t i = 0; i < nout; ++i) c3_out[i] /= tot;
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
            if (q) (void)hipFree(q);
    };
    if ((rc = dev_alloc(&d_at, 3 * (size_t)natoms)) || (rc = dev_alloc(&d_gx, ntheta)) ||
        (rc = dev_alloc(&d_gw, ntheta)) || (rc = dev_alloc(&d_I, npts)) || (rc = dev_alloc(&d_P, P.size())) ||
        (rc = dev_alloc(&d_G, (size_t)KMAX * ntheta * nl1)) || (rc = dev_alloc(&d_c, (size_t)KMAX * LMAX * LMAX)) ||
        (rc = dev_alloc(&d_sean, sean.size()))) {
        cleanup();
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_at, atoms_xyz, sizeof(double) * 3 * natoms, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_gx, gx.data(), sizeof(double) * ntheta, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_gw, gw.data(), sizeof(double) * ntheta, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_P, P.data(), sizeof(double) * P.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sean, sean.data(), sizeof(int) * sean.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_c, 0, sizeof(double2) * (size_t)KMAX * LMAX * LMAX, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_intensity_grid, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, st, d_at, natoms,
                           KMAX, dq, d_gx, ntheta, nphi, d_I);
        const int64_t ng = (int64_t)KMAX * ntheta * nl1;
        hipLaunchKernelGGL(k_intensity_phi, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, st, d_I, KMAX, ntheta,
                           nphi, lfill, d_G);
        const int64_t ncf = (int64_t)KMAX * nl1 * nl1;
        hipLaunchKernelGGL(k_intensity_coeff, dim3((unsigned)((ncf + 255) / 256)), dim3(256), 0, st, d_G, d_P, d_gw,
                           KMAX, ntheta, nphi, lfill, LMAX, d_sean, d_c);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(coeff_out, d_c, sizeof(double2) * (size_t)KMAX * LMAX * LMAX, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    cleanup();
    if (e != hipSuccess) return tp_fail(TP_E_HIP, std::string("intensity kernels: ") + hipGetErrorString(e));
    return 0;
}

// ---- batches ----------------------------------------------------------------
// r must be usable with b's tables for an evaluation with K shells
static int batch_check_ref(const tp_basis *b, const tp_ref *r, int K)
{
    if (!r || !r->b) return tp_fail(TP_E_ARG, "null histogram");
    const tp_basis *rb = r->b;
    if (rb->ctx != b->ctx) return tp_fail(TP_E_STATE, "histogram belongs to another context");
    if (rb->L != b->L || rb->LMAX != b->LMAX || rb->N != b->N)
        return tp_fail(TP_E_ARG, "batch members must share L, LMAX and N");
    if (K > b->K) return tp_fail(TP_E_ARG, "batch member has more shells than the batch basis");
    if (r->ncols > (int64_t)K * (K + 1) * (K + 2) / 6) return tp_fail(TP_E_ARG, "histogram walk longer than K allows");
    return 0;
}

void tp_batch_destroy(tp_batch *B) { batch_free(B); }

int tp_batch_create(tp_ctx *ctx, const tp_basis *b, int ngroups, int nprop, const tp_ref *const *refs,
                    const double *const *coeff, const int64_t *ld, const uint64_t *seeds, int hist_cap,
                    tp_batch **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (int rc = check_ctx(ctx)) return rc;
    if (!b || !refs || !coeff || !ld || !seeds) return tp_fail(TP_E_ARG, "null argument");
    if (b->lam && b->prec != TP_PREC_FP16X3_T64)
        return tp_fail(TP_E_UNSUPPORTED, "batched runs on the curved Ewald sphere (lambda dq != 0) need "
                                         "TP_PREC_FP16X3_T64");
    if (b->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
    if (ngroups < 1 || nprop < 3 || nprop > 4096 || hist_cap < 0)
        return tp_fail(TP_E_ARG, "need ngroups >= 1, 3 <= nprop <= 4096, hist_cap >= 0");
    if (b->L > kRotMaxL || b->L < 2) return tp_fail(TP_E_UNSUPPORTED, "rotation proposals need 2 <= L <= 22");
    std::vector<int> Kg(ngroups);
    for (int g = 0; g < ngroups; ++g) {
        if (!refs[g]) return tp_fail(TP_E_ARG, "null histogram");
        Kg[g] = refs[g]->khi;   // shells that enter the run's energy
        if (int rc = batch_check_ref(b, refs[g], Kg[g])) return rc;
    }
    tp_batch *B = nullptr;
    if (int rc = batch_build(ctx, b, ngroups, nprop, refs, Kg.data(), true, hist_cap, &B)) return rc;
    tp_basis *bm = const_cast<tp_basis *>(b);
    for (int g = 0; g < ngroups; ++g) {
        if (int rc = upload_coeff(ctx, bm, coeff[g], ld[g], B->h_gs[g].K, B->d_cur + (size_t)g * B->cstride,
                                  bm->d_raw)) {
            batch_free(B);
            return rc;
        }
        B->h_gs[g].seed = seeds[g];
    }
    hipError_t e = hipMemcpyAsync(B->d_gs, B->h_gs.data(), sizeof(GroupState) * ngroups, hipMemcpyHostToDevice,
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        batch_free(B);
        return tp_fail(TP_E_HIP, std::string("batch upload: ") + hipGetErrorString(e));
    }
    *out = B;
    return 0;
}

int tp_batch_autotune(tp_batch *B, int reps, double *ms_out, int *config_out)
{
    if (!B || reps < 1) return tp_fail(TP_E_ARG, "null batch or reps < 1");
    if (int rc = check_ctx(B->ctx)) return rc;
    hipStream_t st = B->ctx->stream;
    if (B->sweep) {
        // proposals of the current states (probe stream) so every stage sees real data
        const int nq = B->ne * B->nl;
        hipLaunchKernelGGL(k_quat_random, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, st, B->d_ev, B->ne,
                           B->nl, B->d_gs, 1, B->d_quat);
        hipLaunchKernelGGL(k_rotate_batch, dim3((unsigned)B->nl, (unsigned)B->ne), dim3(256), 0, st, B->d_ev,
                           B->d_quat, B->nl, B->b->n_c, B->d_gs, B->family, 1);
        TP_HIP(hipGetLastError());
    }
    if (int rc = batch_autotune(B, reps)) return rc;
    if (ms_out)
        for (int i = 0; i < 2 + kNumTiles; ++i) ms_out[i] = B->tune_ms[i];
    if (config_out && B->prec == TP_PREC_FP16X3_T64) {
        fac_shape(B->b, B->gcand, &config_out[0], &config_out[1]);
        config_out[2] = B->sym_order;
        config_out[3] = 4 + B->gcand;
    } else if (config_out) {
        config_out[0] = kTiles[B->gcand][0];
        config_out[1] = kTiles[B->gcand][1];
        config_out[2] = B->sym ? B->sym_order : B->order;
        config_out[3] = kTiles[B->gcand][2];
    }
    return 0;
}

int tp_batch_probe(tp_batch *B, double temperature_factor)
{
    if (!B || !B->sweep) return tp_fail(TP_E_ARG, "null batch or not an annealing batch");
    if (int rc = check_ctx(B->ctx)) return rc;
    if (int rc = batch_enqueue_sweep(B, B->ctx->stream, 1, 1.0, 1.0, temperature_factor)) return rc;
    TP_HIP(hipStreamSynchronize(B->ctx->stream));
    return 0;
}

int tp_batch_run(tp_batch *B, int nsweeps, double decay, double factor)
{
    if (!B || !B->sweep || nsweeps < 0) return tp_fail(TP_E_ARG, "null batch, not an annealing batch or nsweeps < 0");
    if (!(decay > 0.0) || !(factor > 0.0)) return tp_fail(TP_E_ARG, "decay and factor must be positive");
    if (int rc = check_ctx(B->ctx)) return rc;
    if (!B->exec || decay != B->g_decay || factor != B->g_factor)
        if (int rc = batch_capture(B, decay, factor)) return rc;
    for (int i = 0; i < nsweeps; ++i) TP_HIP(hipGraphLaunch(B->exec, B->ctx->stream));
    TP_HIP(hipStreamSynchronize(B->ctx->stream));
    return 0;
}

int tp_batch_group(tp_batch *B, int g, double *E, double *T, double *amax, int32_t *nacc, int32_t *nsweep)
{
    if (!B || g < 0 || g >= B->ngroups) return tp_fail(TP_E_ARG, "null batch or group out of range");
    if (int rc = check_ctx(B->ctx)) return rc;
    GroupState s;
    TP_HIP(hipMemcpy(&s, B->d_gs + g, sizeof(GroupState), hipMemcpyDeviceToHost));
    if (E) *E = s.E;
    if (T) *T = s.T;
    if (amax) *amax = s.amax;
    if (nacc) *nacc = s.nacc;
    if (nsweep) *nsweep = s.nsweep;
    return 0;
}

int tp_batch_coeff(tp_batch *B, int g, double *coeff, int64_t ld)
{
    if (!B || !B->sweep || g < 0 || g >= B->ngroups || !coeff) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(B->ctx)) return rc;
    const tp_basis *b = B->b;
    if (ld < (int64_t)b->LMAX * b->LMAX) return tp_fail(TP_E_ARG, "ld shorter than LMAX^2");
    const int K = B->h_gs[g].K;
    std::vector<double2> h((size_t)K * b->n_c);
    TP_HIP(hipMemcpy(h.data(), B->d_cur + (size_t)g * B->cstride, sizeof(double2) * h.size(), hipMemcpyDeviceToHost));
    for (int k = 0; k < K; ++k)
        for (int c = 0; c < b->n_c; ++c) {
            double *d = coeff + 2 * ((int64_t)k * ld + b->sean0[c]);
            d[0] = h[(size_t)k * b->n_c + c].x;
            d[1] = h[(size_t)k * b->n_c + c].y;
        }
    return 0;
}

int tp_batch_history(tp_batch *B, double *Ehist, int n)
{
    if (!B || !Ehist || n < 0) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(B->ctx)) return rc;
    const int m = std::min(n, B->hist_cap);
    if (m > 0)
        TP_HIP(hipMemcpy(Ehist, B->d_Ehist, sizeof(double) * (size_t)m * B->ngroups, hipMemcpyDeviceToHost));
    return 0;
}

int tp_batch_eval_rotations(tp_batch *B, int ngeo, const double *gen, const double *theta, const double *uvb,
                            double *E_out)
{
    if (!B || !B->sweep || !E_out || ngeo < 0 || ngeo > B->nprop) return tp_fail(TP_E_ARG, "bad argument");
    if (ngeo > 0 && (!gen || !theta)) return tp_fail(TP_E_ARG, "ngeo > 0 needs gen and theta");
    if (ngeo < B->nprop && !uvb) return tp_fail(TP_E_ARG, "plane slots need uvb");
    if (int rc = check_ctx(B->ctx)) return rc;
    const int L = B->b->L, nl = B->nl;
    if (L > kRotMaxL) return tp_fail(TP_E_UNSUPPORTED, "given rotations need L <= 22");
    int64_t gstride = 0;
    for (int l = 2; l <= L; l += 2) gstride += (int64_t)(2 * l + 1) * (2 * l + 1);
    const int uvstride = 2 * (2 * L + 1) + 1;
    if (ngeo > 0)
        for (int64_t i = 0; i < (int64_t)B->ne * nl; ++i)
            if (!std::isfinite(theta[i])) return tp_fail(TP_E_ARG, "theta must be finite");
    if (!B->d_gen) {
        int rc;
        if ((rc = dev_alloc(&B->d_gen, (size_t)B->ngroups * gstride)) ||
            (rc = dev_alloc(&B->d_theta, (size_t)B->ne * nl)) ||
            (rc = dev_alloc(&B->d_uvb, (size_t)B->ne * nl * uvstride)))
            return rc;
    }
    hipStream_t st = B->ctx->stream;
    if (ngeo > 0) {
        TP_HIP(hipMemcpyAsync(B->d_gen, gen, sizeof(double) * (size_t)B->ngroups * gstride, hipMemcpyHostToDevice, st));
        TP_HIP(hipMemcpyAsync(B->d_theta, theta, sizeof(double) * (size_t)B->ne * nl, hipMemcpyHostToDevice, st));
    }
    if (ngeo < B->nprop)
        TP_HIP(hipMemcpyAsync(B->d_uvb, uvb, sizeof(double) * (size_t)B->ne * nl * uvstride, hipMemcpyHostToDevice,
                              st));
    hipLaunchKernelGGL(k_rotation_batch, dim3((unsigned)nl, (unsigned)B->ne), dim3(256), 0, st, B->d_ev, nl,
                       B->b->n_c, ngeo, B->d_gen, gstride, B->d_theta, B->d_uvb, uvstride);
    TP_HIP(hipGetLastError());
    if (int rc = batch_enqueue_eval(B, st)) return rc;
    TP_HIP(hipMemcpyAsync(E_out, B->d_E, sizeof(double) * B->ne, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    return 0;
}

int tp_batch_set_state(tp_batch *B, int g, const double *coeff, int64_t ld, double E)
{
    if (!B || !B->sweep || g < 0 || g >= B->ngroups || !coeff) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(B->ctx)) return rc;
    const tp_basis *b = B->b;
    if (ld < (int64_t)b->LMAX * b->LMAX) return tp_fail(TP_E_ARG, "ld shorter than LMAX^2");
    const int K = B->h_gs[g].K;
    std::vector<double2> h((size_t)K * b->n_c);
    for (int k = 0; k < K; ++k)
        for (int c = 0; c < b->n_c; ++c) {
            const double *d = coeff + 2 * ((int64_t)k * ld + b->sean0[c]);
            h[(size_t)k * b->n_c + c] = make_double2(d[0], d[1]);
        }
    // on the batch's stream, after the work already queued there
    hipStream_t st = B->ctx->stream;
    GroupState s;
    TP_HIP(hipMemcpyAsync(&s, B->d_gs + g, sizeof(GroupState), hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    s.E = E;
    TP_HIP(hipMemcpyAsync(B->d_cur + (size_t)g * B->cstride, h.data(), sizeof(double2) * h.size(),
                          hipMemcpyHostToDevice, st));
    TP_HIP(hipMemcpyAsync(B->d_gs + g, &s, sizeof(GroupState), hipMemcpyHostToDevice, st));
    TP_HIP(hipStreamSynchronize(st));
    return 0;
}

int tp_batch_eval_quats(tp_batch *B, const double *quat, double *E_out)
{
    if (!B || !B->sweep || !quat || !E_out) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(B->ctx)) return rc;
    hipStream_t st = B->ctx->stream;
    TP_HIP(hipMemcpyAsync(B->d_quat, quat, sizeof(double) * (size_t)B->ne * B->nl * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_rotate_batch, dim3((unsigned)B->nl, (unsigned)B->ne), dim3(256), 0, st, B->d_ev, B->d_quat,
                       B->nl, B->b->n_c, nullptr, TP_MOVE_QUAT, 0);
    TP_HIP(hipGetLastError());
    if (int rc = batch_enqueue_eval(B, st)) return rc;
    TP_HIP(hipMemcpyAsync(E_out, B->d_E, sizeof(double) * B->ne, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    return 0;
}

int tp_energy_batch(tp_ctx *ctx, int n, const tp_basis *const *bases, const tp_ref *const *refs,
                    const double *const *coeff, const int64_t *ld, double *E_out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (n < 1 || !bases || !refs || !coeff || !ld || !E_out) return tp_fail(TP_E_ARG, "bad argument");
    int im = 0;
    for (int i = 0; i < n; ++i) {
        if (!bases[i] || !refs[i]) return tp_fail(TP_E_ARG, "null basis or histogram");
        if (bases[i]->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
        if (refs[i]->b != bases[i]) return tp_fail(TP_E_STATE, "histogram was built on another basis");
        if (bases[i]->K > bases[im]->K) im = i;
    }
    const tp_basis *bm = bases[im];
    std::vector<int> Kb(n);
    for (int i = 0; i < n; ++i) {
        Kb[i] = bases[i]->K;
        if (int rc = batch_check_ref(bm, refs[i], Kb[i])) return rc;
        if (bases[i]->prec != bm->prec) return tp_fail(TP_E_ARG, "batch members must share the precision mode");
    }
    hipStream_t st = ctx->stream;
    if ((bm->prec != TP_PREC_FP16X3 && bm->prec != TP_PREC_BF16 && bm->prec != TP_PREC_FP16X3_T64) ||
        !bm->use_mfma || (bm->prec != TP_PREC_FP16X3_T64 && bm->R % 128 != 0) ||
        (bm->prec == TP_PREC_FP16X3_T64 && (!bm->use_sym || !(bm->use_sep || (bm->lam && bm->sep_wide))))) {
        // one evaluation after another on the device
        double *dE = nullptr;
        if (int rc = dev_alloc(&dE, n)) return rc;
        int rc = 0;
        for (int i = 0; i < n && !rc; ++i) {
            tp_basis *b = const_cast<tp_basis *>(bases[i]);
            rc = upload_coeff(ctx, b, coeff[i], ld[i], b->K, b->d_coef, b->d_raw);
            if (!rc) rc = enqueue_eval(ctx, b, refs[i], b->d_coef, nullptr, nullptr, 0, nullptr, dE + i, mc_none());
            if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = tp_fail(TP_E_HIP, "evaluation failed");
        }
        if (!rc && hipMemcpy(E_out, dE, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = tp_fail(TP_E_HIP, "hipMemcpy failed");
        (void)hipFree(dE);
        return rc;
    }
    std::vector<const void *> sig;
    sig.push_back(bm);
    for (int i = 0; i < n; ++i) sig.push_back(refs[i]);
    bool fresh = false;
    if (!ctx->eb || ctx->eb_sig != sig) {
        batch_free(ctx->eb);
        ctx->eb = nullptr;
        ctx->eb_sig.clear();
        tp_batch *B = nullptr;
        if (int rc = batch_build(ctx, bm, n, 1, refs, Kb.data(), false, 0, &B)) return rc;
        ctx->eb = B;
        ctx->eb_sig = sig;
        fresh = true;
    }
    tp_batch *B = ctx->eb;
    tp_basis *bmm = const_cast<tp_basis *>(bm);
    for (int i = 0; i < n; ++i)
        if (int rc = upload_coeff(ctx, bmm, coeff[i], ld[i], B->h_ev[i].K, B->d_prop + (size_t)i * B->cstride,
                                  B->d_raw))
            return rc;
    if (fresh) {
        if (int rc = batch_autotune(B, 2)) return rc;
        if (int rc = batch_capture(B, 0.0, 0.0)) return rc;
    }
    TP_HIP(hipGraphLaunch(B->exec, st));
    TP_HIP(hipMemcpyAsync(E_out, B->d_E, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    return 0;
}

int tp_rotate(tp_ctx *ctx, const tp_basis *bc, const double *coeff, int64_t ld, int nshell, const double *quat,
              double *out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!bc || !coeff || !quat || !out || nshell < 1) return tp_fail(TP_E_ARG, "bad argument");
    if (bc->ctx != ctx) return tp_fail(TP_E_STATE, "basis belongs to another context");
    if (bc->L > kRotMaxL) return tp_fail(TP_E_UNSUPPORTED, "rotation needs L <= 22");
    if (ld < (int64_t)bc->LMAX * bc->LMAX) return tp_fail(TP_E_ARG, "ld shorter than LMAX^2");
    tp_basis *b = const_cast<tp_basis *>(bc);
    const int nl = b->L / 2;
    std::memcpy(out, coeff, sizeof(double) * 2 * (size_t)nshell * ld);
    if (nl == 0) return 0;
    double2 *src = nullptr, *dst = nullptr;
    double *dq = nullptr;
    EvalDesc *dev = nullptr;
    int rc = 0;
    if ((rc = dev_alloc(&src, (size_t)nshell * b->n_c)) || (rc = dev_alloc(&dst, (size_t)nshell * b->n_c)) ||
        (rc = dev_alloc(&dq, (size_t)nl * 4)) || (rc = dev_alloc(&dev, 1))) {
        (void)hipFree(src); (void)hipFree(dst); (void)hipFree(dq); (void)hipFree(dev);
        return rc;
    }
    EvalDesc d;
    std::memset(&d, 0, sizeof(d));
    d.src = src;
    d.coef = dst;
    d.K = nshell;
    rc = upload_coeff(ctx, b, coeff, ld, nshell, src, nullptr);
    std::vector<double2> h((size_t)nshell * b->n_c);
    if (!rc) {
        hipError_t e = hipMemcpyAsync(dq, quat, sizeof(double) * nl * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dev, &d, sizeof(d), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_rotate_batch, dim3((unsigned)nl, 1u), dim3(256), 0, ctx->stream, dev, dq, nl, b->n_c,
                               nullptr, TP_MOVE_QUAT, 0);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(h.data(), dst, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = tp_fail(TP_E_HIP, std::string("tp_rotate: ") + hipGetErrorString(e));
    }
    if (!rc)
        for (int k = 0; k < nshell; ++k)
            for (int c = 0; c < b->n_c; ++c) {
                double *o = out + 2 * ((int64_t)k * ld + b->sean0[c]);
                o[0] = h[(size_t)k * b->n_c + c].x;
                o[1] = h[(size_t)k * b->n_c + c].y;
            }
    (void)hipFree(src); (void)hipFree(dst); (void)hipFree(dq); (void)hipFree(dev);
    return rc;
}

// ---- learned proposer -------------------------------------------------------
int tp_mlp_nparam(void) { return kMlpNParam; }

void tp_mlp_destroy(tp_mlp *m) { mlp_free(m); }

int tp_mlp_create(tp_ctx *ctx, uint64_t seed, tp_mlp **out)
{
    if (!out) return tp_fail(TP_E_ARG, "null output pointer");
    *out = nullptr;
    if (int rc = check_ctx(ctx)) return rc;
    tp_mlp *m = new tp_mlp();
    m->ctx = ctx;
    int rc = 0;
    if ((rc = dev_alloc(&m->d_p, kMlpNParam)) || (rc = dev_alloc(&m->d_m, kMlpNParam)) ||
        (rc = dev_alloc(&m->d_v, kMlpNParam)) || (rc = dev_alloc(&m->d_g, kMlpNParam)) || (rc = dev_alloc(&m->d_t, 1))) {
        mlp_free(m);
        return rc;
    }
    for (int L = 0; L < 4; ++L)
        if ((rc = dev_alloc(&m->d_W[L], (size_t)kMlpRows[L] * kMlpCols[L])) ||
            (rc = dev_alloc(&m->d_Wt[L], (size_t)kMlpRows[L] * kMlpCols[L]))) {
            mlp_free(m);
            return rc;
        }
    // He-normal weights (Box-Muller on splitmix64), zero biases; the output
    // layer starts small with bias = identity quaternion (w = 1) per l
    std::vector<float> p(kMlpNParam, 0.f);
    uint64_t st = seed;
    for (int L = 0; L < 4; ++L) {
        const double sd = std::sqrt(2.0 / kMlpCols[L]) * (L == 3 ? 0.1 : 1.0);
        for (int r = 0; r < kMlpRows[L]; ++r)
            for (int c = 0; c < kMlpCols[L]; ++c) {
                const double u1 = ((splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
                const double u2 = ((splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
                const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * tph::kPi * u2);
                if (L == 3 && r >= kMlpQ) continue;
                p[(size_t)kMlpOffW[L] + (size_t)r * kMlpCols[L] + c] = (float)(sd * z);
            }
    }
    for (int li = 0; li < kMlpQ / 4; ++li) p[(size_t)mlp_bias_off(3) + 4 * li] = 1.f;
    if ((rc = mlp_upload(m, p))) {
        mlp_free(m);
        return rc;
    }
    *out = m;
    return 0;
}

int tp_mlp_get_params(tp_mlp *m, float *out)
{
    if (!m || !out) return tp_fail(TP_E_ARG, "null argument");
    if (int rc = check_ctx(m->ctx)) return rc;
    TP_HIP(hipMemcpy(out, m->d_p, sizeof(float) * kMlpNParam, hipMemcpyDeviceToHost));
    return 0;
}

int tp_mlp_set_params(tp_mlp *m, const float *in)
{
    if (!m || !in) return tp_fail(TP_E_ARG, "null argument");
    if (int rc = check_ctx(m->ctx)) return rc;
    std::vector<float> p(in, in + kMlpNParam);
    for (int r = kMlpQ; r < kMlpOut; ++r) {   // output padding stays zero
        for (int c = 0; c < kMlpHid; ++c) p[(size_t)kMlpOffW[3] + (size_t)r * kMlpHid + c] = 0.f;
        p[(size_t)mlp_bias_off(3) + r] = 0.f;
    }
    return mlp_upload(m, p);
}

// host rows [n][192] float -> device BF16 X and X^T (rows padded to a multiple of 64)
static int mlp_put_rows(tp_mlp *m, const float *X, int n, int B)
{
    std::vector<uint16_t> xb((size_t)B * kMlpIn, 0), xt((size_t)B * kMlpIn, 0);
    for (int r = 0; r < n; ++r)
        for (int i = 0; i < kMlpIn; ++i) {
            const uint16_t h = host_bf16(X[(size_t)r * kMlpIn + i]);
            xb[(size_t)r * kMlpIn + i] = h;
            xt[(size_t)i * B + r] = h;
        }
    TP_HIP(hipMemcpy(m->X, xb.data(), sizeof(uint16_t) * xb.size(), hipMemcpyHostToDevice));
    TP_HIP(hipMemcpy(m->Xt, xt.data(), sizeof(uint16_t) * xt.size(), hipMemcpyHostToDevice));
    return 0;
}

int tp_mlp_forward(tp_mlp *m, const float *X, int n, float *Y)
{
    if (!m || !X || !Y || n < 1) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(m->ctx)) return rc;
    const int B = (int)round_up(n, 64);
    if (int rc = mlp_ensure(m, B)) return rc;
    if (int rc = mlp_put_rows(m, X, n, B)) return rc;
    hipStream_t st = m->ctx->stream;
    mlp_forward(m, st, m->X, B, m->H, m->Ht, m->Y);
    TP_HIP(hipGetLastError());
    TP_HIP(hipMemcpyAsync(Y, m->Y, sizeof(float) * (size_t)n * kMlpOut, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    return 0;
}

int tp_mlp_loss_grad(tp_mlp *m, const float *X, const float *T, int n, double *loss, float *grad)
{
    if (!m || !X || !T || n < 1 || n % 64) return tp_fail(TP_E_ARG, "bad argument (n must be a positive multiple of 64)");
    if (int rc = check_ctx(m->ctx)) return rc;
    if (int rc = mlp_ensure(m, n)) return rc;
    if (int rc = mlp_put_rows(m, X, n, n)) return rc;
    hipStream_t st = m->ctx->stream;
    TP_HIP(hipMemcpyAsync(m->T, T, sizeof(float) * (size_t)n * kMlpQ, hipMemcpyHostToDevice, st));
    mlp_forward(m, st, m->X, n, m->H, m->Ht, m->Y);
    mlp_backward(m, st, n);
    TP_HIP(hipGetLastError());
    std::vector<float> lp(n);
    TP_HIP(hipMemcpyAsync(lp.data(), m->lpart, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    if (grad) TP_HIP(hipMemcpyAsync(grad, m->d_g, sizeof(float) * kMlpNParam, hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    if (loss) {
        double s = 0.0;
        for (float v : lp) s += v;
        *loss = s;
    }
    return 0;
}

int tp_mlp_train(tp_mlp *m, const float *X, const float *T, int n, int steps, int batch, float lr, uint64_t seed,
                 float *loss_trace)
{
    if (!m || !X || !T || n < 1 || steps < 0 || batch < 64 || batch % 64)
        return tp_fail(TP_E_ARG, "bad argument (batch must be a positive multiple of 64)");
    if (int rc = check_ctx(m->ctx)) return rc;
    if (n > m->ndata_cap) {
        // a graph captured on the old buffers must not outlive them
        if (m->exec) { (void)hipGraphExecDestroy(m->exec); m->exec = nullptr; }
        if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
        TP_HIP(hipStreamSynchronize(m->ctx->stream));
        if (m->d_Xdata) (void)hipFree(m->d_Xdata);
        if (m->d_Tdata) (void)hipFree(m->d_Tdata);
        m->d_Xdata = nullptr;
        m->d_Tdata = nullptr;
        m->ndata_cap = 0;
        if (int rc = dev_alloc(&m->d_Xdata, (size_t)n * kMlpIn)) return rc;
        if (int rc = dev_alloc(&m->d_Tdata, (size_t)n * kMlpQ)) return rc;
        m->ndata_cap = n;
    }
    if (!m->d_count)
        if (int rc = dev_alloc(&m->d_count, 1)) return rc;
    std::vector<uint16_t> xb((size_t)n * kMlpIn);
    for (size_t i = 0; i < xb.size(); ++i) xb[i] = host_bf16(X[i]);
    TP_HIP(hipMemcpy(m->d_Xdata, xb.data(), sizeof(uint16_t) * xb.size(), hipMemcpyHostToDevice));
    TP_HIP(hipMemcpy(m->d_Tdata, T, sizeof(float) * (size_t)n * kMlpQ, hipMemcpyHostToDevice));
    TP_HIP(hipMemcpy(m->d_count, &n, sizeof(int), hipMemcpyHostToDevice));
    return mlp_train(m, m->d_Xdata, m->d_Tdata, m->d_count, steps, batch, lr, seed, loss_trace);
}

int tp_mlp_train_log(tp_mlp *m, tp_batch *B, int steps, int batch, float lr, uint64_t seed, float *loss_trace)
{
    if (!m || !B || steps < 0 || batch < 64 || batch % 64) return tp_fail(TP_E_ARG, "bad argument");
    if (!B->log_cap) return tp_fail(TP_E_STATE, "the batch keeps no move log (tp_batch_log)");
    if (int rc = check_ctx(m->ctx)) return rc;
    return mlp_train(m, B->d_Xlog, B->d_Tlog, B->d_logn, steps, batch, lr, seed, loss_trace);
}

int tp_batch_log(tp_batch *B, int capacity)
{
    if (!B || !B->sweep || capacity < 1) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(B->ctx)) return rc;
    if (B->log_cap) return tp_fail(TP_E_STATE, "move log already enabled");
    if (int rc = dev_alloc(&B->d_Xlog, (size_t)capacity * kMlpIn)) return rc;
    if (int rc = dev_alloc(&B->d_Tlog, (size_t)capacity * kMlpQ)) return rc;
    if (int rc = dev_alloc(&B->d_logn, 1)) return rc;
    TP_HIP(hipMemset(B->d_logn, 0, sizeof(int)));
    B->log_cap = capacity;
    if (B->exec) return batch_capture(B, B->g_decay, B->g_factor);
    return 0;
}

int tp_batch_log_count(tp_batch *B, int32_t *n)
{
    if (!B || !n) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(B->ctx)) return rc;
    *n = 0;
    if (B->log_cap) TP_HIP(hipMemcpy(n, B->d_logn, sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

int tp_batch_log_read(tp_batch *B, int first, int n, float *X, float *T)
{
    if (!B || !B->log_cap || first < 0 || n < 0 || first + n > B->log_cap) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(B->ctx)) return rc;
    std::vector<uint16_t> xb((size_t)n * kMlpIn);
    TP_HIP(hipMemcpy(xb.data(), B->d_Xlog + (size_t)first * kMlpIn, sizeof(uint16_t) * xb.size(), hipMemcpyDeviceToHost));
    if (T) TP_HIP(hipMemcpy(T, B->d_Tlog + (size_t)first * kMlpQ, sizeof(float) * (size_t)n * kMlpQ, hipMemcpyDeviceToHost));
    if (X)
        for (size_t i = 0; i < xb.size(); ++i) {
            const uint32_t u = (uint32_t)xb[i] << 16;
            float f;
            std::memcpy(&f, &u, 4);
            X[i] = f;
        }
    return 0;
}

int tp_batch_set_proposer(tp_batch *B, tp_mlp *m, int nlearn, double noise)
{
    if (!B || !B->sweep || nlearn < 0 || nlearn >= B->nprop || noise < 0.0) return tp_fail(TP_E_ARG, "bad argument");
    if (nlearn > 0 && !m) return tp_fail(TP_E_ARG, "learned slots need a network");
    if (m && m->ctx != B->ctx) return tp_fail(TP_E_STATE, "network belongs to another context");
    if (int rc = check_ctx(B->ctx)) return rc;
    const int Mi = (int)round_up((int64_t)B->ngroups * std::max(1, nlearn), 64);
    if (Mi > B->Mi) {
        void *ptrs[] = {B->d_Xi, B->d_Hi[0], B->d_Hi[1], B->d_Hi[2], B->d_Yi};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        B->d_Xi = nullptr; B->d_Hi[0] = B->d_Hi[1] = B->d_Hi[2] = nullptr; B->d_Yi = nullptr;
        B->Mi = 0;
        int rc = 0;
        if ((rc = dev_alloc(&B->d_Xi, (size_t)Mi * kMlpIn)) || (rc = dev_alloc(&B->d_Hi[0], (size_t)Mi * kMlpHid)) ||
            (rc = dev_alloc(&B->d_Hi[1], (size_t)Mi * kMlpHid)) || (rc = dev_alloc(&B->d_Hi[2], (size_t)Mi * kMlpHid)) ||
            (rc = dev_alloc(&B->d_Yi, (size_t)Mi * kMlpOut)))
            return rc;
        TP_HIP(hipMemset(B->d_Xi, 0, sizeof(uint16_t) * (size_t)Mi * kMlpIn));
        B->Mi = Mi;
    }
    B->mlp = nlearn > 0 ? m : nullptr;
    B->nlearn = nlearn;
    B->noise = noise;
    std::vector<GroupState> gs(B->ngroups);
    TP_HIP(hipMemcpy(gs.data(), B->d_gs, sizeof(GroupState) * B->ngroups, hipMemcpyDeviceToHost));
    for (auto &g : gs) g.nlearn = nlearn;
    TP_HIP(hipMemcpy(B->d_gs, gs.data(), sizeof(GroupState) * B->ngroups, hipMemcpyHostToDevice));
    for (auto &g : B->h_gs) g.nlearn = nlearn;
    if (B->exec) return batch_capture(B, B->g_decay, B->g_factor);
    return 0;
}

int tp_batch_set_amax(tp_batch *B, double amax)
{
    if (!B || !B->sweep || !(amax > 0.0) || amax > 1.0) return tp_fail(TP_E_ARG, "need 0 < amax <= 1");
    if (int rc = check_ctx(B->ctx)) return rc;
    std::vector<GroupState> gs(B->ngroups);
    TP_HIP(hipMemcpy(gs.data(), B->d_gs, sizeof(GroupState) * B->ngroups, hipMemcpyDeviceToHost));
    for (auto &g : gs) g.amax = amax;
    TP_HIP(hipMemcpy(B->d_gs, gs.data(), sizeof(GroupState) * B->ngroups, hipMemcpyHostToDevice));
    return 0;
}

int tp_batch_set_contraction(tp_batch *B, int cand)
{
    if (!B || cand < 0 || cand >= kNumTiles) return tp_fail(TP_E_ARG, "contraction candidate out of range");
    if (int rc = check_ctx(B->ctx)) return rc;
    if (B->prec == TP_PREC_FP16X3_T64 && cand >= kNumFac)
        return tp_fail(TP_E_ARG, "TP_PREC_FP16X3_T64 batches have contraction candidates 0..3");
    if (B->prec != TP_PREC_FP16X3_T64 && kTiles[cand][2] == 3) {
        const int nrt = (int)(B->b->R_pad16 / kTiles[cand][0]), nct = (int)(B->ldA / kTiles[cand][1]);
        if (nrt % 8 != 0 || (nrt * nct) % 8 != 0)
            return tp_fail(TP_E_UNSUPPORTED, "the XCD-partitioned schedule needs row and tile counts divisible by 8");
    }
    B->gcand = cand;
    TP_HIP(hipStreamSynchronize(B->ctx->stream));
    if (B->exec) return batch_capture(B, B->g_decay, B->g_factor);
    return 0;
}

int tp_batch_set_family(tp_batch *B, int family)
{
    if (!B || !B->sweep || (family != TP_MOVE_QUAT && family != TP_MOVE_PLANE)) return tp_fail(TP_E_ARG, "bad argument");
    if (int rc = check_ctx(B->ctx)) return rc;
    B->family = family;
    if (B->exec) return batch_capture(B, B->g_decay, B->g_factor);
    return 0;
}

}  // extern "C"

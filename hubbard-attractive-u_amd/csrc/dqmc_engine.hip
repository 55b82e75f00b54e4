// dqmc_engine.hip -- host side of the MI355X DQMC sweep engine and its C ABI.
//
// The sweep schedule mirrors DQMC::Core (src/dqmc_core.cpp): the same slice order, the same
// stabilization points (t % pace == 0 || t == nt going up, (t-1) % pace == 0 going down),
// the same left/right stack roles and the same wrap-error bookkeeping.  What differs is how
// each step is computed:
//   * all chains of a context advance together; every step is one batched kernel launch;
//   * B(t) = diag(e^{dt mu + alpha s_t}) exp(dt nnt A) is applied through fused diagonal
//     scalings of one shared exp(dt nnt A) (the chemical potential commutes with the hopping
//     matrix, so a per-chain mu costs nothing);
//   * the UDV stacks are built by Householder QR on columns pre-sorted by norm instead of
//     SVD, and the stacks keep the accumulated V (no V_0 V_1 ... recomputation);
//   * (1 + L R^T)^{-1} is recomposed by solving with a QR factorization instead of an
//     explicit LU inverse.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/dqmc_amd.h"
#include "dqmc_kernels.h"

using namespace dqmc;

namespace {

struct DStack {
    double* U = nullptr;  // [B][S][N*N]
    double* V = nullptr;  // [B][S][N*N]  accumulated V
    double* D = nullptr;  // [B][S][N]
    int len = 0;
};

const uint64_t kMinstdM = 2147483647ULL;

double host_canonical(uint64_t& x)
{
    x = (x * 16807ULL) % kMinstdM;
    const uint64_t x1 = x;
    x = (x * 16807ULL) % kMinstdM;
    const uint64_t x2 = x;
    const double R = 2147483646.0;
    const double R2 = 4611686009837453312.0;
    double u = std::fma((double)(x2 - 1ULL), R, (double)(x1 - 1ULL)) / R2;
    if (u >= 1.0) u = std::nextafter(1.0, 0.0);
    return u;
}

}  // namespace

#define MSPLIT_MAX 4
struct dqmc_ctx {
    // parameters
    int nl = 0, N = 0, L = 0, pace = 0, S = 0, B = 0, prec = DQMC_FP64, kmax = 1;
    double dt = 0, beta = 0, nnt = 0, u = 0, mu = 0, alpha = 0;
    bool eqtime = false, dynamic = false, thermalization = false, use_graphs = false;
    bool initialized = false;
    // QR of the stabilizations: 0 = blocked (panel kernel + MFMA trailing GEMMs), 1 = one-kernel
    // column-panel QR (DQMC_QR=v1)
    int qr_mode = 0;
    // slice products exp(-+dt K) X as Kronecker squares of the ring exponential (8 MFMAs per
    // lattice vector) instead of dense N x N GEMMs (DQMC_WRAP=dense selects the GEMMs)
    bool kron = true;
    bool trsm_v1 = false;   // DQMC_TRSM=v1: one-kernel thread-per-column back-substitution
    // diagonal block of the blocked back-substitution (DQMC_TRSM_BLOCK; 0: 64 rows from N = 64 up, else
    // 32; at N = 256, 256 chains: 348 us with 64, 402 with 32, 468 with 128 for 256 right-hand sides)
    int trsm_block = 0;
    // the recompose system's two products and right-hand side in one launch, and the stabilization's
    // compare-and-commit in the epilogue of the last product (DQMC_RECOMPOSE_FUSE=0: two GEMMs and a
    // copy, and the compare-and-commit kernel)
    bool recompose_fuse = true;
    int qr_update = QR_UPDATE_REG;   // QR trailing update (DQMC_QR_UPDATE=gemm: three GEMMs)
    int qr_panel = QR_PANEL_256;     // QR panel: 4 waves per chain (DQMC_QR_PANEL=1024: 16 waves)
    bool solve_lu = true;            // recompose solve by partial-pivot LU (DQMC_SOLVE=qr: Householder QR)
    int* lu_perm = nullptr;          // [B][N] row order of the LU panel rows
    int metro_threads = 1024;  // Metropolis workgroup size at N = 256 (DQMC_METRO_THREADS)
    double *qrV = nullptr, *qrT = nullptr, *qrWt = nullptr;   // blocked-QR scratch: V, T, Wt and Wt T
    // look-ahead of the blocked QR / LU (DQMC_LOOKAHEAD=0 turns it off): second stream, fork/join
    // events and the parity-1 panel buffers
    bool lookahead = true;
    hipStream_t stream2 = nullptr;
    hipEvent_t la_fork = nullptr, la_join = nullptr;
    double *qrV2 = nullptr, *qrT2 = nullptr;
    int* lu_perm2 = nullptr;
    LookAhead la() const
    {
        LookAhead a;
        if (!lookahead) return a;
        a.s2 = stream2; a.fork = la_fork; a.join = la_join; a.V2 = qrV2; a.T2 = qrT2; a.P2 = lu_perm2;
        a.fuse = panel_fuse && !capturing;
        return a;
    }
    int current_t = 0;
    long long updating_sweeps = 0;
    std::vector<double> chain_mu;
    std::vector<uint64_t> rng_host;
    std::string err;
    int tile64 = TILE_64x64, tile32 = TILE_64x64;
    float tune_ms[2][TILE_COUNT] = {};      // autotuner timings per precision and tile

    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    // per-launch kernel timer (benchmark instrumentation)
    int kt_kernel = -1, kt_used = 0;
    std::vector<hipEvent_t> kt_ev;
    std::map<int, hipGraphExec_t> graphs;

    // device memory
    std::vector<void*> allocs;
    double *expK = nullptr, *invExpK = nullptr;
    float *expK32 = nullptr, *invExpK32 = nullptr;
    double *E1 = nullptr, *F1 = nullptr;       // exp(+-dt nnt A1) of the ring, 16x16 (nl <= 16) or 32x32 padded
    double *kS1 = nullptr, *kS2 = nullptr;     // transposes for the right product above 16 x 16
    float *E1f = nullptr, *F1f = nullptr;
    double *tab = nullptr, *itab = nullptr, *delta = nullptr;
    uint64_t *powtab = nullptr, *rng = nullptr;
    unsigned long long* acc_count = nullptr;
    uint32_t* accmask = nullptr;
    int8_t* fields = nullptr;
    double *G64 = nullptr;
    float *G32 = nullptr, *T32 = nullptr;
    double *stkA = nullptr, *stkB = nullptr, *stk = nullptr, *stk_alt = nullptr;
    double* ident = nullptr;               // one N x N identity, read at batch stride 0
    DStack* chain_stack = nullptr;         // the stack the current B-chain block is pushed onto
    bool chain_first = false;              // no factor of the block applied yet
    // B-chain factors on the 16 x 16 lattice are collected and applied in one pass at the push
    // (DQMC_CHAIN=slice applies each one when its slice is done)
    bool chain_fused = true;
    int chain_pend[KCHAIN_MAX] = {};
    int chain_npend = 0;
    bool chain_pend_adj = false;
    // FP64 16 x 16 sweeps: the Metropolis workgroups of a slice also apply the wrap that follows it
    // (up sweeps: to the next slice unless a stabilization comes between; down sweeps: the slice's
    // own); DQMC_WRAP_FUSE=0 turns it off
    bool wrap_fuse = true;
    bool wrap_done = false;
    // Eager sweeps: a look-ahead panel applies the previous update to its own columns (one launch and
    // one kernel boundary fewer per panel; DQMC_PANEL_FUSE=0 turns it off).  In a captured graph both
    // branches of a look-ahead step start together, and the combined launch, which then runs its update
    // part beside the whole rest of the update, measured slower than a separate launch for those
    // columns (3415 vs 3485 sweeps/s); eagerly the panel is dispatched first (3499 vs 3417).
    bool panel_fuse = true;
    bool capturing = false;
    // Metropolis launches of groups of the chains on their own streams between the points where all
    // chains meet again (metro_join); DQMC_METRO_SPLIT=n sets the group count (1: all chains at once)
    int metro_groups = 2;
    bool split_open = false;
    hipStream_t ms_stream[MSPLIT_MAX] = {};   // [0] is the main stream, [1] the second one
    hipEvent_t ms_fork = nullptr, ms_join[MSPLIT_MAX] = {};
    double *W1 = nullptr, *W2 = nullptr, *W3 = nullptr, *qr = nullptr;
    int* perm = nullptr;
    double *gt0A = nullptr, *gt0B = nullptr, *g0tA = nullptr, *g0tB = nullptr, *gt0 = nullptr, *g0t = nullptr;
    double *vecgftt = nullptr, *vecgft0 = nullptr, *vecgf0t = nullptr;
    double *wrap_eq = nullptr, *wrap_dyn = nullptr, *maxdG = nullptr;
    uint32_t* health = nullptr;      // [B] HEALTH_* flags of the current call
    // observables measured on the device: cache [B][cstride] (equal-time scalars, G(k,t) [nk][L],
    // LDOS [L], chi_perp [L]), bins [B][nbins][cstride], per-slice partials, cos(k . r_d) [nk][N]
    unsigned obs_mask = 0;
    int nbins = 0, nk = 0, cstride = 0;
    long long eq_count = 0, dyn_count = 0;
    double *meas_cache = nullptr, *meas_data = nullptr, *meas_part = nullptr, *costab = nullptr;
    double* wd_bound = nullptr;      // watchdog bound on max|G_recomputed - G_working| (0 = off)
    double* cmp_part = nullptr;      // [B][CMP_PART_SLOTS][3] partial reductions of compare_commit
    double* colnorm = nullptr;       // [B][N] squared column norms of a pushed block's chain (16 x 16)
    bool colnorm_ready = false;      // colnorm holds those of the chain in stk
    bool norm_fuse = true;           // DQMC_NORM_FUSE=0: the pre-pivot reads the chain for its norms
    bool rhs_identity = true;        // DQMC_RHS_IDENTITY=0: the pre-pivot writes the push QR's identity
    DStack left, right;
    // parameter adjoint at fixed fields (allocated on first use)
    int* nb = nullptr;
    int8_t* fields_save = nullptr;
    double *adj_partial = nullptr, *adj_grad = nullptr, *adj_gdiag = nullptr, *adj_glogdet = nullptr,
           *adj_out = nullptr;

    long long NN() const { return (long long)N * N; }
    // Row-major recompose / UDV systems [A | rhs] are N rows of weq() (equal-time: N + N) or wdyn()
    // (time-displaced: N + 2N) doubles.
    int weq() const { return 2 * N; }
    int wdyn() const { return 3 * N; }
    long long sqr() const { return (long long)N * wdyn(); }

    int fail(int code, const char* fmt, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    template <typename P>
    int alloc(P** p, size_t bytes)
    {
        void* q = nullptr;
        if (hipMalloc(&q, bytes < 16 ? 16 : bytes) != hipSuccess) return fail(DQMC_ENOMEM, "hipMalloc(%zu) failed", bytes);
        allocs.push_back(q);
        *p = (P*)q;
        return 0;
    }
};

#define HCHECK(c, expr)                                                                            \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess) return (c)->fail(DQMC_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                                               __FILE__, __LINE__);                                \
    } while (0)
#define RCHECK(expr)                \
    do {                            \
        int _r = (expr);            \
        if (_r) return _r;          \
    } while (0)

namespace dqmc {
namespace {
thread_local std::string g_free_err;
}
int free_error(int code, const char* msg)
{
    g_free_err = msg;
    return code;
}
const char* free_error_text() { return g_free_err.c_str(); }
}  // namespace dqmc

namespace {

enum { KT_METRO = 0, KT_QR = 1, KT_GEMM = 2, KT_TRSM = 3, KT_KRON = 4 };
int kt_pre(dqmc_ctx* c, int id)
{
    if (c->kt_kernel != id || c->kt_used + 2 > (int)c->kt_ev.size()) return 0;
    HCHECK(c, hipEventRecord(c->kt_ev[c->kt_used], c->stream));
    return 0;
}
int kt_post(dqmc_ctx* c, int id)
{
    if (c->kt_kernel != id || c->kt_used + 2 > (int)c->kt_ev.size()) return 0;
    HCHECK(c, hipEventRecord(c->kt_ev[c->kt_used + 1], c->stream));
    c->kt_used += 2;
    return 0;
}
#define TIMED(c, id, call)          \
    do {                            \
        RCHECK(kt_pre((c), (id)));  \
        HCHECK((c), (call));        \
        RCHECK(kt_post((c), (id))); \
    } while (0)

// exp(s * dt * nnt * A) for the periodic nl x nl square lattice, from its plane-wave
// eigenbasis: the adjacency (src/square_lattice.cpp:48-61) has eigenvalues
// 2 cos kx + 2 cos ky with multiplicities counted like the accumulated entries.
void lattice_expk(int nl, double x, std::vector<double>& out)
{
    const int N = nl * nl;
    if (x == 0.0) {   // nnt = 0: exactly the identity (the cosine sums below round to ~1e-17)
        out.assign((size_t)N * N, 0.0);
        for (int i = 0; i < N; ++i) out[i + (size_t)i * N] = 1.0;
        return;
    }
    std::vector<double> f(N, 0.0);  // value for displacement (dx, dy)
    for (int dy = 0; dy < nl; ++dy)
        for (int dx = 0; dx < nl; ++dx) {
            double s = 0.0;
            for (int my = 0; my < nl; ++my)
                for (int mx = 0; mx < nl; ++mx) {
                    const double kx = 2.0 * M_PI * mx / nl, ky = 2.0 * M_PI * my / nl;
                    s += std::cos(kx * dx + ky * dy) * std::exp(x * 2.0 * (std::cos(kx) + std::cos(ky)));
                }
            f[dx + nl * dy] = s / N;
        }
    out.assign((size_t)N * N, 0.0);
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            const int dx = ((j % nl) - (i % nl) + nl) % nl, dy = ((j / nl) - (i / nl) + nl) % nl;
            out[i + (size_t)j * N] = f[dx + nl * dy];
        }
}

DiagArg dnone() { return DiagArg(); }
DiagArg dvec(int mode, const double* v, long long vstride)
{
    DiagArg d;
    d.mode = mode; d.v = v; d.vstride = vstride;
    return d;
}
DiagArg dfield(dqmc_ctx* c, const double* table, int t)
{
    DiagArg d;
    d.mode = DIAG_FIELD; d.v = table; d.vstride = 2;
    d.f = c->fields + (long long)t * c->N; d.fstride = (long long)c->L * c->N;
    return d;
}

template <typename T>
int gemm(dqmc_ctx* c, const T* A, long long sA, int lda, int tA, const T* B, long long sB, int ldb, int tB, T* C,
         long long sC, int ldc, int tC, DiagArg rowd = DiagArg(), DiagArg innd = DiagArg(),
         DiagArg cold = DiagArg(), double alpha = 1.0, double beta = 0.0)
{
    GemmArgs<T> g;
    g.M = c->N; g.Nc = c->N; g.K = c->N;
    g.A = A; g.sA = sA; g.lda = lda; g.tA = tA;
    g.B = B; g.sB = sB; g.ldb = ldb; g.tB = tB;
    g.C = C; g.sC = sC; g.ldc = ldc; g.tC = tC;
    g.alpha = alpha; g.beta = beta;
    g.rowd = rowd; g.innd = innd; g.cold = cold;
    g.batch = c->B;
    TIMED(c, KT_GEMM, launch_gemm<T>(g, sizeof(T) == 8 ? c->tile64 : c->tile32, c->stream));
    return 0;
}

// Whether the push's QR takes its right-hand side [N, 2N) as the identity without reading it (the blocked
// QR with the fused update at N <= 256), so that the pre-pivot need not write it.
bool qr_takes_identity(const dqmc_ctx* c)
{
    return c->rhs_identity && c->qr_mode != 1 && c->qr_update == QR_UPDATE_REG && c->N <= 256;
}
hipError_t qr_any(dqmc_ctx* c, double* buf, long long stride, int N, int W, int batch, hipStream_t s,
                  bool rhs_identity = false)
{
    if (c->qr_mode == 1) return launch_qr_v1(buf, stride, N, W, batch, s);
    const LookAhead la = c->la();
    return launch_qr_blocked(buf, stride, N, W, batch, c->qrV, c->qrT, c->qrWt, (long long)c->wdyn() * 32, c->tile64, c->qr_update,
                             c->qr_panel, s, &la, rhs_identity);
}
hipError_t trsm_any(dqmc_ctx* c, double* buf, long long stride, int N, int W, int c0, int P, int batch, hipStream_t s)
{
    if (c->trsm_v1) return launch_trsm_upper_v1(buf, stride, N, W, c0, P, batch, s);
    const int block = c->trsm_block ? c->trsm_block : (N >= 64 ? 64 : 32);
    return launch_trsm_blocked(buf, stride, N, W, c0, P, batch, c->tile64, s, block);
}

// Factorization of the recompose system [temp | rhs] (row-major N x W in c->qr): partial-pivot LU
// like the reference's PartialPivLU (stable_numerics.hpp:235, 304, 322), or Householder QR.  Either
// leaves an upper triangle for the back-substitution and the transformed right-hand sides.
hipError_t factor_any(dqmc_ctx* c, int W)
{
    if (c->solve_lu)
    {
        const LookAhead la = c->la();
        return launch_lu_blocked(c->qr, c->sqr(), c->N, W, c->B, c->qrV, c->qrT, c->lu_perm, c->health, c->stream, &la);
    }
    return qr_any(c, c->qr, c->sqr(), c->N, W, c->B, c->stream);
}

// C = diag(r) (E(x)E) diag(s) X diag(cv)  (left) or  diag(r) X diag(s) (E(x)E) diag(cv)  (right)
template <typename T>
int kron(dqmc_ctx* c, bool left, const T* X, T* C, const T* E, DiagArg rowd = DiagArg(), DiagArg innd = DiagArg(),
         DiagArg cold = DiagArg(), long long sX = -1)
{
    KronArgs<T> k;
    k.X = X; k.sX = sX < 0 ? c->NN() : sX; k.C = C; k.sC = c->NN(); k.E = E;
    k.nl = c->nl; k.N = c->N;
    k.rowd = rowd; k.innd = innd; k.cold = cold;
    k.batch = c->B;
    if (c->nl > 16) {
        // 17..32 sites per side: FP64 only (the engine keeps dense products for MIXED there)
        if constexpr (std::is_same<T, double>::value) {
            if (left) TIMED(c, KT_KRON, launch_kron32_left(k, c->stream));
            else TIMED(c, KT_KRON, launch_kron32_right(k, c->kS1, c->kS2, c->stream));
            return 0;
        } else {
            return c->fail(DQMC_ESTATE, "kron: FP32 products need nl <= 16");
        }
    }
    if (left) TIMED(c, KT_KRON, launch_kron_left<T>(k, c->stream));
    else TIMED(c, KT_KRON, launch_kron_right<T>(k, c->stream));
    return 0;
}

// ---- slice operations --------------------------------------------------------------------
// All chains meet again: the main stream waits for the second stream's Metropolis launches.
int metro_join(dqmc_ctx* c)
{
    if (!c->split_open) return 0;
    for (int g = 1; g < c->metro_groups; ++g) {
        HCHECK(c, hipEventRecord(c->ms_join[g], c->ms_stream[g]));
        HCHECK(c, hipStreamWaitEvent(c->stream, c->ms_join[g], 0));
    }
    c->split_open = false;
    return 0;
}
// G <- B(t) G B(t)^{-1}  (Core::wrap_from_0_to_beta, src/dqmc_core.cpp:142-147)
int wrap_up(dqmc_ctx* c, int t)
{
    if (!c->wrap_done) RCHECK(metro_join(c));
    const long long nn = c->NN();
    const int N = c->N;
    if (c->kron) {
        // D (E(x)E) G (F(x)F) D^{-1}: the chemical potential sits in the D tables
        if (c->prec == DQMC_MIXED) {
            RCHECK(kron<float>(c, true, c->G32, c->T32, c->E1f, dfield(c, c->tab, t)));
            RCHECK(kron<float>(c, false, c->T32, c->G32, c->F1f, dnone(), dnone(), dfield(c, c->itab, t)));
        } else {
            // the previous slice's Metropolis workgroups may have applied this wrap already
            if (c->wrap_done) {
                c->wrap_done = false;
                return 0;
            }
            RCHECK(kron<double>(c, true, c->G64, c->W1, c->E1, dfield(c, c->tab, t)));
            RCHECK(kron<double>(c, false, c->W1, c->G64, c->F1, dnone(), dnone(), dfield(c, c->itab, t)));
        }
        return 0;
    }
    if (c->prec == DQMC_MIXED) {
        RCHECK(gemm<float>(c, c->expK32, 0, N, 0, c->G32, nn, N, 0, c->T32, nn, N, 0));
        RCHECK(gemm<float>(c, c->T32, nn, N, 0, c->invExpK32, 0, N, 0, c->G32, nn, N, 0, dfield(c, c->tab, t), dnone(),
                           dfield(c, c->itab, t)));
    } else {
        RCHECK(gemm<double>(c, c->expK, 0, N, 0, c->G64, nn, N, 0, c->W1, nn, N, 0));
        RCHECK(gemm<double>(c, c->W1, nn, N, 0, c->invExpK, 0, N, 0, c->G64, nn, N, 0, dfield(c, c->tab, t), dnone(),
                            dfield(c, c->itab, t)));
    }
    return 0;
}
// G <- B(t)^{-1} G B(t)  (Core::wrap_from_beta_to_0, src/dqmc_core.cpp:156-161)
int wrap_down(dqmc_ctx* c, int t)
{
    if (!c->wrap_done) RCHECK(metro_join(c));
    const long long nn = c->NN();
    const int N = c->N;
    if (c->kron) {
        // (F(x)F) D^{-1} G D (E(x)E)
        if (c->prec == DQMC_MIXED) {
            RCHECK(kron<float>(c, false, c->G32, c->T32, c->E1f, dnone(), dfield(c, c->tab, t)));
            RCHECK(kron<float>(c, true, c->T32, c->G32, c->F1f, dnone(), dfield(c, c->itab, t)));
        } else if (c->wrap_done) {
            // applied by the slice's Metropolis workgroups, left product first (the reference applies
            // B from the right first; the two orders agree to rounding)
            c->wrap_done = false;
        } else {
            RCHECK(kron<double>(c, false, c->G64, c->W1, c->E1, dnone(), dfield(c, c->tab, t)));
            RCHECK(kron<double>(c, true, c->W1, c->G64, c->F1, dnone(), dfield(c, c->itab, t)));
        }
        return 0;
    }
    if (c->prec == DQMC_MIXED) {
        RCHECK(gemm<float>(c, c->G32, nn, N, 0, c->expK32, 0, N, 0, c->T32, nn, N, 0, dnone(), dfield(c, c->tab, t)));
        RCHECK(gemm<float>(c, c->invExpK32, 0, N, 0, c->T32, nn, N, 0, c->G32, nn, N, 0, dnone(), dfield(c, c->itab, t)));
    } else {
        RCHECK(gemm<double>(c, c->G64, nn, N, 0, c->expK, 0, N, 0, c->W1, nn, N, 0, dnone(), dfield(c, c->tab, t)));
        RCHECK(gemm<double>(c, c->invExpK, 0, N, 0, c->W1, nn, N, 0, c->G64, nn, N, 0, dnone(), dfield(c, c->itab, t)));
    }
    return 0;
}
// X_out = B(t) X (rowd = D) or B(t)^T X (innd = D) with one shared exp(-dt K)
int left_expk(dqmc_ctx* c, const double* X, double* C, const double* table, int t, bool inner)
{
    const long long nn = c->NN();
    const DiagArg d = dfield(c, table, t);
    if (c->kron) return kron<double>(c, true, X, C, table == c->tab ? c->E1 : c->F1, inner ? dnone() : d, inner ? d : dnone());
    return gemm<double>(c, table == c->tab ? c->expK : c->invExpK, 0, c->N, 0, X, nn, c->N, 0, C, nn, c->N, 0,
                        inner ? dnone() : d, inner ? d : dnone());
}
// X_out = X exp(+-dt K) diag(d) (col scaling) or X diag(d) exp(+-dt K) (inner scaling)
int right_expk(dqmc_ctx* c, const double* X, double* C, const double* table, int t, bool inner)
{
    const long long nn = c->NN();
    const DiagArg d = dfield(c, table, t);
    const bool fwd = table == c->tab;   // tab goes with exp(-dt K), itab with exp(+dt K)
    if (c->kron)
        return kron<double>(c, false, X, C, fwd ? c->E1 : c->F1, dnone(), inner ? d : dnone(), inner ? dnone() : d);
    return gemm<double>(c, X, nn, c->N, 0, fwd ? c->expK : c->invExpK, 0, c->N, 0, C, nn, c->N, 0, dnone(),
                        inner ? d : dnone(), inner ? dnone() : d);
}
// Whether the Metropolis workgroups can also make the next wrap's left product (FP64 Kronecker
// products on the 16 x 16 lattice).
bool wrap_fusable(dqmc_ctx* c)
{
    return c->wrap_fuse && c->kron && c->prec == DQMC_FP64 && c->nl == 16;
}
// Metropolis update of slice t.  With a fused wrap the workgroups then apply the next wrap as well
// (the left product into W1, the right one back into G), and that wrap call does nothing: FUSE_UP is
// wrap_up(t') with t' = lt, G <- D(t') (E(x)E) G (F(x)F) D(t')^{-1}; FUSE_DOWN is wrap_down(t),
// G <- ((F(x)F) D(t)^{-1} G) D(t) (E(x)E).
enum WrapFuse { FUSE_NONE = 0, FUSE_UP = 1, FUSE_DOWN = 2 };
int metropolis(dqmc_ctx* c, int t, int fuse = FUSE_NONE, int lt = 0)
{
    if (c->prec == DQMC_MIXED) {
        MetroArgs<float> a;
        a.G = c->G32; a.sG = c->NN();
        a.fields = c->fields; a.N = c->N; a.L = c->L; a.t = t;
        a.rng = c->rng; a.powtab = c->powtab; a.delta = c->delta;
        a.accmask = c->accmask; a.acc_count = c->acc_count; a.health = c->health; a.kmax = c->kmax;
        TIMED(c, KT_METRO, launch_metropolis<float>(a, c->B, c->metro_threads, c->stream));
    } else {
        MetroArgs<double> a;
        a.G = c->G64; a.sG = c->NN();
        a.fields = c->fields; a.N = c->N; a.L = c->L; a.t = t;
        a.rng = c->rng; a.powtab = c->powtab; a.delta = c->delta;
        a.accmask = c->accmask; a.acc_count = c->acc_count; a.health = c->health; a.kmax = c->kmax;
        if (fuse == FUSE_UP) {
            a.lout = c->W1; a.lE = c->E1; a.ltab = c->tab; a.lt = lt; a.lmode = 1;
            a.rE = c->F1; a.rtab = c->itab; a.rmode = 1;
        } else if (fuse == FUSE_DOWN) {
            a.lout = c->W1; a.lE = c->F1; a.ltab = c->itab; a.lt = t; a.lmode = 2;
            a.rE = c->E1; a.rtab = c->tab; a.rmode = 2;
        }
        const int ng = c->metro_groups;
        if (ng > 1 && c->B >= 8 * ng && c->kt_kernel < 0) {
            // The groups run on their own streams, each in slice order; the chains are independent, so
            // a group starts its next slice as soon as its own slowest chain is done, and fills the CUs
            // the other groups' tails leave idle.
            if (!c->split_open) {
                HCHECK(c, hipEventRecord(c->ms_fork, c->stream));
                for (int g = 1; g < ng; ++g) HCHECK(c, hipStreamWaitEvent(c->ms_stream[g], c->ms_fork, 0));
                c->split_open = true;
            }
            for (int g = 0; g < ng; ++g) {
                const int b0 = (int)((long long)c->B * g / ng), b1 = (int)((long long)c->B * (g + 1) / ng);
                a.chain0 = b0;
                HCHECK(c, launch_metropolis<double>(a, b1 - b0, c->metro_threads, c->ms_stream[g]));
            }
        } else {
            TIMED(c, KT_METRO, launch_metropolis<double>(a, c->B, c->metro_threads, c->stream));
        }
        c->wrap_done = fuse != FUSE_NONE;
    }
    return 0;
}

// ---- UDV stack push (svd_stack.hpp:107-175 restated with QR) ------------------------------
double* sU(dqmc_ctx* c, DStack& s, int k) { return s.U + (long long)k * c->NN(); }
double* sV(dqmc_ctx* c, DStack& s, int k) { return s.V + (long long)k * c->NN(); }
double* sD(dqmc_ctx* c, DStack& s, int k) { return s.D + (long long)k * c->N; }

// Start the B-chain of a stack block that will be pushed onto s.
void chain_begin(dqmc_ctx* c, DStack& s)
{
    c->stk = c->stkA;
    c->stk_alt = c->stkB;
    c->chain_stack = &s;
    c->chain_first = true;
    c->chain_npend = 0;
    c->colnorm_ready = false;
}
// One factor of the block's B-chain: stk <- B(t) stk (src/dqmc_core.cpp:190) or B(t)^T stk (:251, adjB,
// src/hubbard.cpp:333-341).  The first factor of a block reads the top U of the stack the block goes to
// (the identity when that stack is empty) instead of an identity-initialized chain, and the last one
// (`last`) scales the columns by that stack's top D, so the push receives (chain * U_top) * diag(D_top)
// (svd_stack.hpp:144) with no separate N x N x N product.
// The operand of the block's next factor: the chain so far, or the stack top's U (the identity at
// batch stride 0 for an empty stack) before the first factor.
const double* chain_operand(dqmc_ctx* c, long long& sX)
{
    DStack& s = *c->chain_stack;
    const long long nn = c->NN();
    sX = nn;
    if (!c->chain_first) return c->stk;
    if (s.len > 0) { sX = (long long)c->S * nn; return sU(c, s, s.len - 1); }
    sX = 0;
    return c->ident;
}
// Apply the collected factors of the block in one pass (16 x 16 lattice, FP64 chain).
int chain_flush(dqmc_ctx* c, bool last)
{
    if (c->chain_npend == 0) return 0;
    RCHECK(metro_join(c));   // the factors read the slices' fields
    DStack& s = *c->chain_stack;
    KronChainArgs k;
    k.X = chain_operand(c, k.sX);
    k.C = c->stk_alt; k.sC = c->NN();
    k.E = c->E1;
    k.fields = c->fields; k.fstride = (long long)c->L * c->N;
    k.table = c->tab;
    for (int i = 0; i < c->chain_npend; ++i) k.ts[i] = c->chain_pend[i];
    k.n = c->chain_npend;
    k.adjoint = c->chain_pend_adj ? 1 : 0;
    k.cold = (last && s.len > 0) ? dvec(DIAG_VEC, sD(c, s, s.len - 1), (long long)c->S * c->N) : dnone();
    k.batch = c->B;
    // the block's last pass also sums its columns' squared norms for the push's pre-pivot
    k.norms = (last && c->norm_fuse) ? c->colnorm : nullptr;
    c->colnorm_ready = last && c->norm_fuse;
    TIMED(c, KT_KRON, launch_kron_chain16(k, c->stream));
    std::swap(c->stk, c->stk_alt);
    c->chain_first = false;
    c->chain_npend = 0;
    return 0;
}
int chain_mul(dqmc_ctx* c, int t, bool adjoint, bool last)
{
    DStack& s = *c->chain_stack;
    const long long nn = c->NN();
    if (c->chain_fused && c->kron && c->nl == 16) {
        // the factors of a block are applied when it is pushed, after its last slice's update
        if (c->chain_npend > 0 && c->chain_pend_adj != adjoint) RCHECK(chain_flush(c, false));
        c->chain_pend[c->chain_npend++] = t;
        c->chain_pend_adj = adjoint;
        if (last || c->chain_npend == KCHAIN_MAX) RCHECK(chain_flush(c, last));
        return 0;
    }
    RCHECK(metro_join(c));   // the factor reads the slice's fields
    long long sX = nn;
    const double* X = chain_operand(c, sX);
    const DiagArg d = dfield(c, c->tab, t);
    const DiagArg rowd = adjoint ? dnone() : d, innd = adjoint ? d : dnone();
    const DiagArg cold = (last && s.len > 0) ? dvec(DIAG_VEC, sD(c, s, s.len - 1), (long long)c->S * c->N) : dnone();
    if (c->kron) RCHECK(kron<double>(c, true, X, c->stk_alt, c->E1, rowd, innd, cold, sX));
    else RCHECK(gemm<double>(c, c->expK, 0, c->N, 0, X, sX, c->N, 0, c->stk_alt, nn, c->N, 0, rowd, innd, cold));
    std::swap(c->stk, c->stk_alt);
    c->chain_first = false;
    return 0;
}

int stack_push(dqmc_ctx* c, DStack& s)
{
    const long long nn = c->NN(), st = (long long)c->S * nn, sd = (long long)c->S * c->N;
    const int N = c->N;
    if (s.len >= c->S) return c->fail(DQMC_ESTATE, "stack overflow");
    if (c->chain_stack != &s || c->chain_first) return c->fail(DQMC_ESTATE, "stack push without its B-chain");
    // M = (stacks * U_top) * diag(D_top) (svd_stack.hpp:144), formed by the block's chain_mul calls
    const double* M = c->stk;
    const bool ident = qr_takes_identity(c);
    HCHECK(c, launch_udv_prepare(M, nn, c->qr, c->sqr(), c->weq(), c->perm, N, c->B, c->stream,
                                 c->colnorm_ready ? c->colnorm : nullptr, !ident));
    c->colnorm_ready = false;
    TIMED(c, KT_QR, qr_any(c, c->qr, c->sqr(), N, c->weq(), c->B, c->stream, ident));
    HCHECK(c, launch_udv_finish(c->qr, c->sqr(), c->weq(), c->perm, sU(c, s, s.len), st, sD(c, s, s.len), sd, c->W2, nn, N,
                                c->health, c->B, c->stream));
    if (s.len == 0)
        HCHECK(c, launch_copy_d(c->W2, nn, sV(c, s, 0), st, nn, c->B, c->stream));
    else
        RCHECK(gemm<double>(c, sV(c, s, s.len - 1), st, N, 0, c->W2, nn, N, 0, sV(c, s, s.len), st, N, 0));
    s.len += 1;
    return 0;
}

// (1 + u diag(d) v^T)^{-1} with u orthogonal (compute_gf_00bb, stable_numerics.hpp:136-155);
// result written to out, transposed when `transposed_result` (left-stack-empty case).
int solve_00bb(dqmc_ctx* c, const double* u, const double* v, const double* d, double* out, bool transposed_result)
{
    const long long nn = c->NN(), st = (long long)c->S * nn, sd = (long long)c->S * c->N;
    const int N = c->N;
    const int W = c->weq();
    HCHECK(c, launch_build_00bb(u, st, v, st, d, sd, c->qr, c->sqr(), N, W, c->B, c->stream));
    TIMED(c, KT_QR, qr_any(c, c->qr, c->sqr(), N, W, c->B, c->stream));
    TIMED(c, KT_TRSM, trsm_any(c, c->qr, c->sqr(), N, W, N, N, c->B, c->stream));
    HCHECK(c, launch_extract(c->qr, c->sqr(), W, N, out, nn, N, transposed_result ? 0 : 1, c->B, c->stream));
    return 0;
}

// The recompose system in qr (row-major N x W): columns [0, N) hold
//   temp = Dlmax^{-1} (Ul^T Ur) Drmax^{-1} + Dlmin (Vl^T Vr) Drmin   (stable_numerics.hpp:220-233)
// and columns [N, 2N) the right-hand side Dlmax^{-1} Ul^T.  With the 128 x 128 FP64 tile both products
// are one launch over K = 2N that also stores the right-hand side from its staged operand; otherwise
// two GEMMs and a copy.
int recompose_system(dqmc_ctx* c, const double* Ul, const double* Vl, const double* Dl, const double* Ur,
                     const double* Vr, const double* Dr, int W)
{
    const long long st = (long long)c->S * c->NN(), sd = (long long)c->S * c->N;
    const int N = c->N;
    if (c->recompose_fuse && c->tile64 == TILE_128x128 && N >= 128) {
        GemmArgs<double> g;
        g.M = N; g.Nc = N; g.K = N;
        g.A = Ul; g.sA = st; g.lda = N; g.tA = 1;
        g.B = Ur; g.sB = st; g.ldb = N; g.tB = 0;
        g.C = c->qr; g.sC = c->sqr(); g.ldc = W; g.tC = 1;
        g.alpha = 1.0; g.beta = 0.0;
        g.rowd = dvec(DIAG_INVMAX1, Dl, sd); g.cold = dvec(DIAG_INVMAX1, Dr, sd);
        g.batch = c->B;
        GemmPair p;
        p.A2 = Vl; p.sA2 = st; p.lda2 = N;
        p.B2 = Vr; p.sB2 = st; p.ldb2 = N;
        p.rowd2 = dvec(DIAG_MIN1, Dl, sd); p.cold2 = dvec(DIAG_MIN1, Dr, sd);
        p.R = c->qr + N; p.sR = c->sqr(); p.ldr = W;
        TIMED(c, KT_GEMM, launch_gemm_pair(g, p, c->stream));
        return 0;
    }
    RCHECK(gemm<double>(c, Ul, st, N, 1, Ur, st, N, 0, c->qr, c->sqr(), W, 1, dvec(DIAG_INVMAX1, Dl, sd), dnone(),
                        dvec(DIAG_INVMAX1, Dr, sd)));
    RCHECK(gemm<double>(c, Vl, st, N, 1, Vr, st, N, 0, c->qr, c->sqr(), W, 1, dvec(DIAG_MIN1, Dl, sd), dnone(),
                        dvec(DIAG_MIN1, Dr, sd), 1.0, 1.0));
    HCHECK(c, launch_build_rhs(Ul, st, dvec(DIAG_INVMAX1, Dl, sd), 1.0, c->qr, c->sqr(), N, W, N, c->B, c->stream));
    return 0;
}

// Equal-time G from the stack tops into `out` (compute_equaltime_gf, stable_numerics.hpp:186-239).
// With `committed` set, the general case may instead compare the new G with the working G64 and commit it
// in the epilogue of its last product (the wrap error and health folded as launch_compare_commit does);
// *committed then says so, and `out` is not written.
int recompute_eq(dqmc_ctx* c, double* out, bool* committed = nullptr)
{
    const long long nn = c->NN(), st = (long long)c->S * nn, sd = (long long)c->S * c->N;
    const int N = c->N, W = c->weq();
    DStack &Ls = c->left, &Rs = c->right;
    if (Ls.len == 0) {
        // G^T = (1 + Ur Dr Vr^T)^{-1}
        return solve_00bb(c, sU(c, Rs, Rs.len - 1), sV(c, Rs, Rs.len - 1), sD(c, Rs, Rs.len - 1), out, true);
    }
    if (Rs.len == 0) return solve_00bb(c, sU(c, Ls, Ls.len - 1), sV(c, Ls, Ls.len - 1), sD(c, Ls, Ls.len - 1), out, false);
    const double *Ul = sU(c, Ls, Ls.len - 1), *Vl = sV(c, Ls, Ls.len - 1), *Dl = sD(c, Ls, Ls.len - 1);
    const double *Ur = sU(c, Rs, Rs.len - 1), *Vr = sV(c, Rs, Rs.len - 1), *Dr = sD(c, Rs, Rs.len - 1);
    // temp = Dlmax^{-1} (Ul^T Ur) Drmax^{-1} + Dlmin (Vl^T Vr) Drmin, written row-major into qr, and the
    // right-hand side Dlmax^{-1} Ul^T beside it
    RCHECK(recompose_system(c, Ul, Vl, Dl, Ur, Vr, Dr, W));
    TIMED(c, KT_QR, factor_any(c, W));
    TIMED(c, KT_TRSM, trsm_any(c, c->qr, c->sqr(), N, W, N, N, c->B, c->stream));
    // G = Ur Drmax^{-1} X
    if (committed && c->recompose_fuse && c->prec != DQMC_MIXED && c->tile64 == TILE_128x128 && N >= 128 &&
        N <= 256) {
        GemmArgs<double> g;
        g.M = N; g.Nc = N; g.K = N;
        g.A = Ur; g.sA = st; g.lda = N; g.tA = 0;
        g.B = c->qr + N; g.sB = c->sqr(); g.ldb = W; g.tB = 1;
        g.C = c->G64; g.sC = nn; g.ldc = N; g.tC = 0;
        g.alpha = 1.0; g.beta = 0.0;
        g.innd = dvec(DIAG_INVMAX1, Dr, sd);
        g.batch = c->B;
        TIMED(c, KT_GEMM, launch_gemm_commit(g, c->cmp_part, CMP_PART_SLOTS, c->stream));
        const int tiles = ((N + 127) / 128) * ((N + 127) / 128);
        HCHECK(c, launch_compare_fold(c->cmp_part, tiles, c->wrap_eq, c->maxdG, c->wd_bound, c->health, c->B,
                                      c->stream));
        *committed = true;
        return 0;
    }
    RCHECK(gemm<double>(c, Ur, st, N, 0, c->qr + N, c->sqr(), W, 1, out, nn, N, 0, dnone(), dvec(DIAG_INVMAX1, Dr, sd)));
    return 0;
}

// Dynamic recompute (compute_dynamic_gf, stable_numerics.hpp:245-324) into gtt, gt0, g0t.
int recompute_dyn(dqmc_ctx* c, double* gtt, double* gt0n, double* g0tn)
{
    const long long nn = c->NN(), st = (long long)c->S * nn, sd = (long long)c->S * c->N;
    const int N = c->N;
    DStack &Ls = c->left, &Rs = c->right;
    if (Ls.len == 0) {
        RCHECK(solve_00bb(c, sU(c, Rs, Rs.len - 1), sV(c, Rs, Rs.len - 1), sD(c, Rs, Rs.len - 1), gtt, true));
        HCHECK(c, launch_copy_d(gtt, nn, gt0n, nn, nn, c->B, c->stream));
        HCHECK(c, launch_axpi(gtt, nn, g0tn, nn, 1.0, -1.0, N, c->B, c->stream));
        return 0;
    }
    if (Rs.len == 0) {
        RCHECK(solve_00bb(c, sU(c, Ls, Ls.len - 1), sV(c, Ls, Ls.len - 1), sD(c, Ls, Ls.len - 1), gtt, false));
        HCHECK(c, launch_axpi(gtt, nn, gt0n, nn, -1.0, 1.0, N, c->B, c->stream));
        HCHECK(c, launch_axpi(gtt, nn, g0tn, nn, -1.0, 0.0, N, c->B, c->stream));
        return 0;
    }
    const double *Ul = sU(c, Ls, Ls.len - 1), *Vl = sV(c, Ls, Ls.len - 1), *Dl = sD(c, Ls, Ls.len - 1);
    const double *Ur = sU(c, Rs, Rs.len - 1), *Vr = sV(c, Rs, Rs.len - 1), *Dr = sD(c, Rs, Rs.len - 1);
    const int W = c->wdyn();
    RCHECK(recompose_system(c, Ul, Vl, Dl, Ur, Vr, Dr, W));
    HCHECK(c, launch_build_rhs(Vl, st, dvec(DIAG_MIN1, Dl, sd), 1.0, c->qr, c->sqr(), N, W, 2 * N, c->B, c->stream));
    TIMED(c, KT_QR, factor_any(c, W));
    TIMED(c, KT_TRSM, trsm_any(c, c->qr, c->sqr(), N, W, N, 2 * N, c->B, c->stream));
    RCHECK(gemm<double>(c, Ur, st, N, 0, c->qr + N, c->sqr(), W, 1, gtt, nn, N, 0, dnone(), dvec(DIAG_INVMAX1, Dr, sd)));
    RCHECK(gemm<double>(c, Ur, st, N, 0, c->qr + 2 * N, c->sqr(), W, 1, gt0n, nn, N, 0, dnone(),
                        dvec(DIAG_INVMAX1, Dr, sd)));
    // G(0,t) = -B(beta,t) G(t,t) = -(Vr Dr Ur^T) Ur Drmax^{-1} X1 = -Vr Drmin X1, with
    // X1 = temp^{-1} Dlmax^{-1} Ul^T from the solve above.  This needs only Ur orthogonal;
    // the reference's form (stable_numerics.hpp:305-323) also assumes an orthogonal Vl, which
    // holds for its SVD stacks but not for QR-based ones.
    RCHECK(gemm<double>(c, Vr, st, N, 0, c->qr + N, c->sqr(), W, 1, g0tn, nn, N, 0, dnone(), dvec(DIAG_MIN1, Dr, sd),
                        dnone(), -1.0));
    return 0;
}

int store_current(dqmc_ctx* c, double* vec, int t)
{
    RCHECK(metro_join(c));
    const long long nn = c->NN(), sv = (long long)c->L * nn;
    if (c->prec == DQMC_MIXED)
        HCHECK(c, launch_copy_f2d(c->G32, nn, vec + (long long)t * nn, sv, nn, c->B, c->stream));
    else
        HCHECK(c, launch_copy_d(c->G64, nn, vec + (long long)t * nn, sv, nn, c->B, c->stream));
    return 0;
}

int stabilize_eq(dqmc_ctx* c)
{
    RCHECK(metro_join(c));
    bool committed = false;
    RCHECK(recompute_eq(c, c->W3, &committed));
    if (!committed)
        HCHECK(c, launch_compare_commit(c->W3, c->G64, c->prec == DQMC_MIXED ? c->G32 : nullptr, c->NN(), c->N,
                                        c->wrap_eq, c->maxdG, c->wd_bound, c->health, c->cmp_part, c->B, c->stream));
    return 0;
}

// ---- sweeps ---------------------------------------------------------------------------------
int do_sweep_up(dqmc_ctx* c)
{
    const int L = c->L;
    const bool store = !c->thermalization && c->eqtime;
    const bool fuse = wrap_fusable(c);
    c->wrap_done = false;
    chain_begin(c, c->left);
    for (int t = 1; t <= L; ++t) {
        const bool push = t % c->pace == 0 || t == L;
        RCHECK(wrap_up(c, t % L));
        // unless a stabilization follows, the next step is wrap_up(t + 1); when measuring, G(t) is
        // stored first
        RCHECK(metropolis(c, t % L, (fuse && !push && !store) ? FUSE_UP : FUSE_NONE, (t + 1) % L));
        RCHECK(chain_mul(c, t % L, false, push));
        if (push) {
            c->right.len -= 1;
            RCHECK(stack_push(c, c->left));
            RCHECK(stabilize_eq(c));
            chain_begin(c, c->left);
        }
        if (store) RCHECK(store_current(c, c->vecgftt, t % L));
    }
    return metro_join(c);
}

int do_sweep_down(dqmc_ctx* c)
{
    const int L = c->L;
    const bool store = !c->thermalization && c->eqtime;
    const bool fuse = wrap_fusable(c);
    c->wrap_done = false;
    chain_begin(c, c->right);
    for (int t = L; t >= 1; --t) {
        const bool push = (t - 1) % c->pace == 0;
        // the next step on G is always wrap_down(t)
        RCHECK(metropolis(c, t % L, fuse ? FUSE_DOWN : FUSE_NONE));
        RCHECK(chain_mul(c, t % L, true, push));
        RCHECK(wrap_down(c, t % L));
        if (push) {
            c->left.len -= 1;
            RCHECK(stack_push(c, c->right));
            RCHECK(stabilize_eq(c));
            chain_begin(c, c->right);
        }
        if (store) RCHECK(store_current(c, c->vecgftt, (t - 1) % L));
    }
    return metro_join(c);
}

int do_sweep_dynamic(dqmc_ctx* c)
{
    const int L = c->L, N = c->N;
    const long long nn = c->NN(), sv = (long long)L * nn;
    c->gt0 = c->gt0A; c->g0t = c->g0tA;
    c->wrap_done = false;
    if (c->prec == DQMC_MIXED) HCHECK(c, launch_copy_f2d(c->G32, nn, c->W1, nn, nn, c->B, c->stream));
    else HCHECK(c, launch_copy_d(c->G64, nn, c->W1, nn, nn, c->B, c->stream));
    HCHECK(c, launch_copy_d(c->W1, nn, c->gt0, nn, nn, c->B, c->stream));
    HCHECK(c, launch_axpi(c->W1, nn, c->g0t, nn, 1.0, -1.0, N, c->B, c->stream));
    chain_begin(c, c->left);
    for (int t = 1; t <= L; ++t) {
        const int ti = t % L;
        const bool push = t % c->pace == 0 || t == L;
        RCHECK(wrap_up(c, ti));
        double* gt0n = (c->gt0 == c->gt0A) ? c->gt0B : c->gt0A;
        double* g0tn = (c->g0t == c->g0tA) ? c->g0tB : c->g0tA;
        RCHECK(left_expk(c, c->gt0, gt0n, c->tab, ti, false));
        RCHECK(right_expk(c, c->g0t, g0tn, c->itab, ti, false));
        c->gt0 = gt0n; c->g0t = g0tn;
        RCHECK(chain_mul(c, ti, false, push));
        if (push) {
            c->right.len -= 1;
            RCHECK(stack_push(c, c->left));
            RCHECK(recompute_dyn(c, c->W3, c->W2, c->W1));
            HCHECK(c, launch_compare_commit(c->W3, c->G64, c->prec == DQMC_MIXED ? c->G32 : nullptr, nn, N, c->wrap_eq,
                                            c->maxdG, c->wd_bound, c->health, c->cmp_part, c->B, c->stream));
            HCHECK(c, launch_compare_only(c->W2, c->gt0, nn, N, c->wrap_dyn, c->health, c->B, c->stream));
            HCHECK(c, launch_compare_only(c->W1, c->g0t, nn, N, c->wrap_dyn, c->health, c->B, c->stream));
            HCHECK(c, launch_copy_d(c->W2, nn, c->gt0, nn, nn, c->B, c->stream));
            HCHECK(c, launch_copy_d(c->W1, nn, c->g0t, nn, nn, c->B, c->stream));
            chain_begin(c, c->left);
        }
        RCHECK(store_current(c, c->vecgftt, ti));
        if (ti != 0) {
            HCHECK(c, launch_copy_d(c->gt0, nn, c->vecgft0 + (long long)ti * nn, sv, nn, c->B, c->stream));
            HCHECK(c, launch_copy_d(c->g0t, nn, c->vecgf0t + (long long)ti * nn, sv, nn, c->B, c->stream));
        } else {
            RCHECK(store_current(c, c->vecgft0, 0));
            HCHECK(c, launch_axpi(c->vecgft0, sv, c->vecgf0t, sv, 1.0, -1.0, N, c->B, c->stream));
        }
    }
    return 0;
}

// Core::initialize_svdstacks + initialize_green_functions (src/dqmc_core.cpp:67-97) from the
// current fields: right stack = B-chain blocks from beta down to 0, G = G(0).  On return the
// QR buffer still holds the factored (1 + Ur Dr Vr^T) system of the full chain.
int build_from_fields(dqmc_ctx* c)
{
    const int L = c->L;
    const long long nn = c->NN();
    c->left.len = 0; c->right.len = 0; c->current_t = 0;
    chain_begin(c, c->right);
    for (int t = L; t >= 1; --t) {
        const bool push = (t - 1) % c->pace == 0;
        RCHECK(chain_mul(c, t % L, true, push));
        if (push) {
            RCHECK(stack_push(c, c->right));
            chain_begin(c, c->right);
        }
    }
    RCHECK(recompute_eq(c, c->W3));
    HCHECK(c, launch_finite_check(c->W3, nn, c->N, c->health, c->B, c->stream));
    HCHECK(c, launch_copy_d(c->W3, nn, c->G64, nn, nn, c->B, c->stream));
    if (c->prec == DQMC_MIXED) HCHECK(c, launch_copy_d2f(c->W3, nn, c->G32, nn, nn, c->B, c->stream));
    return 0;
}

// Numeric health of the chains over one ABI call: cleared before the call's work, read after it.
int health_reset(dqmc_ctx* c)
{
    HCHECK(c, launch_fill_u32(c->health, c->B, 0u, c->stream));
    return 0;
}
const char* health_text(uint32_t h)
{
    if (h & HEALTH_NONFINITE_G) return "non-finite Green's function at a stabilization";
    if (h & HEALTH_ZERO_PIVOT) return "zero or non-finite pivot in a UDV factorization";
    if (h & HEALTH_RATIO) return "non-finite acceptance ratio";
    return "wrap error above the watchdog bound";
}
int health_check(dqmc_ctx* c, const char* what)
{
    std::vector<uint32_t> h(c->B);
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(h.data(), c->health, (size_t)c->B * 4, hipMemcpyDeviceToHost));
    int bad = 0, first = -1;
    for (int b = 0; b < c->B; ++b)
        if (h[b]) { ++bad; if (first < 0) first = b; }
    if (!bad) return 0;
    return c->fail(DQMC_ENUMERIC, "%s: chain %d: %s (flags 0x%x; %d chain(s) flagged)", what, first,
                   health_text(h[first]), h[first], bad);
}

// Buffers of the fixed-field forward and the adjoint: the dynamic G vectors and the
// neighbour table of the hopping matrix (square_lattice.cpp:30-61: directions x+1, y+1, x-1, y-1).
int ensure_adjoint_buffers(dqmc_ctx* c)
{
    const long long nn = c->NN();
    const size_t B = c->B, L = c->L, N = c->N;
    if (!c->vecgftt) RCHECK(c->alloc(&c->vecgftt, B * L * nn * 8));
    if (!c->vecgft0) {
        RCHECK(c->alloc(&c->vecgft0, B * L * nn * 8)); RCHECK(c->alloc(&c->vecgf0t, B * L * nn * 8));
        RCHECK(c->alloc(&c->gt0A, B * nn * 8)); RCHECK(c->alloc(&c->gt0B, B * nn * 8));
        RCHECK(c->alloc(&c->g0tA, B * nn * 8)); RCHECK(c->alloc(&c->g0tB, B * nn * 8));
    }
    if (!c->nb) {
        RCHECK(c->alloc(&c->nb, N * 4 * sizeof(int)));
        RCHECK(c->alloc(&c->fields_save, B * L * N));
        RCHECK(c->alloc(&c->adj_partial, B * (L + 1) * 3 * 8));
        RCHECK(c->alloc(&c->adj_grad, B * 3 * 8));
        RCHECK(c->alloc(&c->adj_gdiag, B * L * N * 8));
        RCHECK(c->alloc(&c->adj_glogdet, B * 8));
        RCHECK(c->alloc(&c->adj_out, B * 8));
        std::vector<int> nb(N * 4);
        const int nl = c->nl;
        for (int i = 0; i < (int)N; ++i) {
            const int x = i % nl, y = i / nl;
            nb[4 * i + 0] = (x + 1) % nl + nl * y;
            nb[4 * i + 1] = x + nl * ((y + 1) % nl);
            nb[4 * i + 2] = (x - 1 + nl) % nl + nl * y;
            nb[4 * i + 3] = x + nl * ((y - 1 + nl) % nl);
        }
        HCHECK(c, hipMemcpy(c->nb, nb.data(), nb.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    return 0;
}

// The run statistics are a property of the Markov chain; the auxiliary passes of the forward
// and the adjoint leave them as they were.
struct SavedStats {
    std::vector<double> we, wd, dg;
};
int save_stats(dqmc_ctx* c, SavedStats& s)
{
    s.we.resize(c->B); s.wd.resize(c->B); s.dg.resize(c->B);
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(s.we.data(), c->wrap_eq, c->B * 8, hipMemcpyDeviceToHost));
    HCHECK(c, hipMemcpy(s.wd.data(), c->wrap_dyn, c->B * 8, hipMemcpyDeviceToHost));
    HCHECK(c, hipMemcpy(s.dg.data(), c->maxdG, c->B * 8, hipMemcpyDeviceToHost));
    return 0;
}
int restore_stats(dqmc_ctx* c, const SavedStats& s)
{
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(c->wrap_eq, s.we.data(), c->B * 8, hipMemcpyHostToDevice));
    HCHECK(c, hipMemcpy(c->wrap_dyn, s.wd.data(), c->B * 8, hipMemcpyHostToDevice));
    HCHECK(c, hipMemcpy(c->maxdG, s.dg.data(), c->B * 8, hipMemcpyHostToDevice));
    return 0;
}

// Run a sweep eagerly, or through a captured graph keyed by its kind and store flag.
typedef int (*sweep_fn)(dqmc_ctx*);
int run_sweep(dqmc_ctx* c, sweep_fn fn, int kind)
{
    if (!c->use_graphs || c->kt_kernel >= 0) {
        RCHECK(fn(c));
        HCHECK(c, hipStreamSynchronize(c->stream));
        return 0;
    }
    const int key = kind * 2 + ((!c->thermalization && c->eqtime) ? 1 : 0);
    auto it = c->graphs.find(key);
    const int lenL = c->left.len, lenR = c->right.len;
    if (it == c->graphs.end()) {
        hipGraph_t g;
        HCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        int rc = fn(c);
        c->capturing = false;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) return rc;
        HCHECK(c, e);
        hipGraphExec_t ex;
        HCHECK(c, hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        HCHECK(c, hipGraphDestroy(g));
        c->graphs[key] = ex;
        it = c->graphs.find(key);
        // the capture only recorded the launches; restore the host-side state and replay
        c->left.len = lenL; c->right.len = lenR;
    }
    HCHECK(c, hipGraphLaunch(it->second, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    // host-side bookkeeping of the replayed sweep
    if (kind == 1) { c->right.len = c->S; c->left.len = 0; }
    else { c->left.len = c->S; c->right.len = 0; }
    return 0;
}

int autotune(dqmc_ctx* c)
{
    // Pick the GEMM tile per precision by timing the wrap GEMM shape on this context's batch.
    hipEvent_t e0, e1;
    HCHECK(c, hipEventCreate(&e0));
    HCHECK(c, hipEventCreate(&e1));
    const long long nn = c->NN();
    HCHECK(c, launch_set_identity(c->W1, nn, c->N, 1.0, c->B, c->stream));
    if (c->prec == DQMC_MIXED) HCHECK(c, launch_copy_d2f(c->W1, nn, c->T32, nn, nn, c->B, c->stream));
    for (int p = 0; p < 2; ++p) {
        if (p == 1 && c->prec != DQMC_MIXED) break;
        float best = 1e30f;
        int bt = TILE_64x64;
        for (int tile = 0; tile < TILE_COUNT; ++tile) {
            if (p == 0) c->tile64 = tile; else c->tile32 = tile;
            for (int rep = 0; rep < 2; ++rep) {
                HCHECK(c, hipEventRecord(e0, c->stream));
                for (int it = 0; it < 3; ++it) {
                    if (p == 0) RCHECK(gemm<double>(c, c->expK, 0, c->N, 0, c->W1, nn, c->N, 0, c->W2, nn, c->N, 0));
                    else RCHECK(gemm<float>(c, c->expK32, 0, c->N, 0, c->T32, nn, c->N, 0, (float*)c->W2, nn, c->N, 0));
                }
                HCHECK(c, hipEventRecord(e1, c->stream));
                HCHECK(c, hipEventSynchronize(e1));
                float ms = 0;
                HCHECK(c, hipEventElapsedTime(&ms, e0, e1));
                if (rep == 1) c->tune_ms[p][tile] = ms / 3;
                if (rep == 1 && ms < best) { best = ms; bt = tile; }
            }
        }
        if (p == 0) c->tile64 = bt; else c->tile32 = bt;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

}  // namespace

// ================================== C ABI ====================================================
extern "C" {

int dqmc_create(const dqmc_params_c* p, int n_chains, int precision, int delay_k, dqmc_ctx** out)
{
    using dqmc::free_error;
    if (!p || !out || n_chains <= 0) return free_error(DQMC_EINVAL, "dqmc_create: null argument or n_chains < 1");
    *out = nullptr;
    if (p->nl < 1 || p->nt < 1 || p->stabilization_pace < 1 || p->u < 0.0 || p->beta < 0.0)
        return free_error(DQMC_EINVAL, "dqmc_create: need nl >= 1, nt >= 1, stabilization_pace >= 1, u >= 0, beta >= 0");
    if (precision != DQMC_FP64 && precision != DQMC_MIXED) return free_error(DQMC_EINVAL, "dqmc_create: bad precision");
    const int N = p->nl * p->nl;
    if (p->nl > 32) return free_error(DQMC_EINVAL, "dqmc_create: nl > 32 is not supported (N <= 1024)");
    if (delay_k < 1 || delay_k > 32) return free_error(DQMC_EINVAL, "dqmc_create: delay_k must be 1..32");
    const size_t esz = precision == DQMC_MIXED ? 4 : 8;
    // The pending rank-k update of a site block lives in LDS: on lattices above 16 x 16 the delay is
    // capped at what fits (the accept/reject sequence does not depend on it).
    while (delay_k > 1 && metropolis_lds_bytes(N, delay_k, esz) > 160 * 1024) --delay_k;
    if (metropolis_lds_bytes(N, delay_k, esz) > 160 * 1024)
        return free_error(DQMC_EINVAL, "dqmc_create: the Metropolis block does not fit in LDS");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return free_error(DQMC_ENODEV, "dqmc_create: no GPU");

    dqmc_ctx* c = new dqmc_ctx();
    c->nl = p->nl; c->N = N; c->L = p->nt; c->pace = p->stabilization_pace;
    c->S = (c->L % c->pace == 0) ? c->L / c->pace : c->L / c->pace + 1;
    c->B = n_chains; c->prec = precision; c->kmax = delay_k;
    c->beta = p->beta; c->dt = p->beta / p->nt; c->nnt = p->nnt; c->u = p->u; c->mu = p->mu;
    c->alpha = std::acosh(std::exp(0.5 * c->dt * c->u));   // src/hubbard.cpp:35
    c->eqtime = p->collect_eqtime != 0; c->dynamic = p->collect_dynamic != 0;
    {
        const char* q = std::getenv("DQMC_QR");
        c->qr_mode = (q && std::string(q) == "v1") ? 1 : 0;
        const char* sv = std::getenv("DQMC_SOLVE");
        c->solve_lu = !(sv && std::string(sv) == "qr");
        const char* qp = std::getenv("DQMC_QR_PANEL");
        if (qp && std::string(qp) == "1024") c->qr_panel = QR_PANEL_1024;
        const char* qu = std::getenv("DQMC_QR_UPDATE");
        if (qu && std::string(qu) == "gemm") c->qr_update = QR_UPDATE_GEMM;
        const char* lk = std::getenv("DQMC_LOOKAHEAD");
        if (lk && std::string(lk) == "0") c->lookahead = false;
        const char* mt = std::getenv("DQMC_METRO_THREADS");
        if (mt) {
            const int v = std::atoi(mt);
            // the pipelined flush needs waves beside the site-loop wave: 256, 512 or 1024 threads
            if (v != 256 && v != 512 && v != 1024) {
                delete c;
                return free_error(DQMC_EINVAL, "dqmc_create: DQMC_METRO_THREADS must be 256, 512 or 1024");
            }
            c->metro_threads = v;
        }
        const char* tr = std::getenv("DQMC_TRSM");
        c->trsm_v1 = tr && std::string(tr) == "v1";
        const char* tb = std::getenv("DQMC_TRSM_BLOCK");
        if (tb) {
            const int v = std::atoi(tb);
            if (v != 32 && v != 64 && v != 128) {
                delete c;
                return free_error(DQMC_EINVAL, "dqmc_create: DQMC_TRSM_BLOCK must be 32, 64 or 128");
            }
            c->trsm_block = v;
        }
        const char* w = std::getenv("DQMC_WRAP");
        c->kron = !(w && std::string(w) == "dense");
        const char* ri = std::getenv("DQMC_RHS_IDENTITY");
        c->rhs_identity = !(ri && std::string(ri) == "0");
        const char* nf = std::getenv("DQMC_NORM_FUSE");
        c->norm_fuse = !(nf && std::string(nf) == "0");
        const char* rf = std::getenv("DQMC_RECOMPOSE_FUSE");
        c->recompose_fuse = !(rf && std::string(rf) == "0");
        const char* pf = std::getenv("DQMC_PANEL_FUSE");
        c->panel_fuse = !(pf && std::string(pf) == "0");
        const char* msp = std::getenv("DQMC_METRO_SPLIT");
        if (msp) {
            const int v = std::atoi(msp);
            if (v != 0 && v != 1 && v != 2 && v != 4) {
                delete c;
                return free_error(DQMC_EINVAL, "dqmc_create: DQMC_METRO_SPLIT must be 1, 2 or 4");
            }
            c->metro_groups = v < 1 ? 1 : v;
        }
        const char* wf = std::getenv("DQMC_WRAP_FUSE");
        c->wrap_fuse = !(wf && std::string(wf) == "0");
        const char* ch = std::getenv("DQMC_CHAIN");
        c->chain_fused = !(ch && std::string(ch) == "slice");
        if (N > 256) {
            // Lattices above 16 x 16: factored slice products on the 32 x 32 padded ring factor in
            // FP64 (dense GEMMs for MIXED), the Householder QR solve with the wide panel kernel and
            // the three-GEMM update
            if (precision == DQMC_MIXED) c->kron = false;
            c->solve_lu = false;
            c->qr_update = QR_UPDATE_GEMM;
            c->lookahead = false;
            if (c->qr_mode == 1) {
                delete c;
                return free_error(DQMC_EINVAL, "dqmc_create: DQMC_QR=v1 supports N <= 256 only");
            }
        }
    }
    c->chain_mu.assign(n_chains, p->mu);
    c->rng_host.assign(n_chains, 1ULL);
    *out = c;

    const int B = n_chains, L = c->L, S = c->S;
    const long long nn = c->NN();
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return c->fail(DQMC_EHIP, "stream");
    if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) return c->fail(DQMC_EHIP, "stream");
    HCHECK(c, hipEventCreateWithFlags(&c->la_fork, hipEventDisableTiming));
    HCHECK(c, hipEventCreateWithFlags(&c->la_join, hipEventDisableTiming));
    HCHECK(c, hipEventCreateWithFlags(&c->ms_fork, hipEventDisableTiming));
    c->ms_stream[0] = c->stream;
    c->ms_stream[1] = c->stream2;
    for (int g = 2; g < c->metro_groups; ++g)
        if (hipStreamCreateWithFlags(&c->ms_stream[g], hipStreamNonBlocking) != hipSuccess) return c->fail(DQMC_EHIP, "stream");
    for (int g = 1; g < c->metro_groups; ++g) HCHECK(c, hipEventCreateWithFlags(&c->ms_join[g], hipEventDisableTiming));
    HCHECK(c, prepare_kernels());
    for (auto& e : c->ev) HCHECK(c, hipEventCreate(&e));
    RCHECK(c->alloc(&c->expK, nn * 8)); RCHECK(c->alloc(&c->invExpK, nn * 8));
    RCHECK(c->alloc(&c->E1, 1024 * 8)); RCHECK(c->alloc(&c->F1, 1024 * 8));
    if (c->kron && N > 256) { RCHECK(c->alloc(&c->kS1, (size_t)B * nn * 8)); RCHECK(c->alloc(&c->kS2, (size_t)B * nn * 8)); }
    RCHECK(c->alloc(&c->ident, nn * 8));
    RCHECK(c->alloc(&c->E1f, 256 * 4)); RCHECK(c->alloc(&c->F1f, 256 * 4));
    RCHECK(c->alloc(&c->tab, (size_t)B * 16)); RCHECK(c->alloc(&c->itab, (size_t)B * 16));
    RCHECK(c->alloc(&c->delta, (size_t)B * 16));
    RCHECK(c->alloc(&c->powtab, (size_t)2 * N * 8)); RCHECK(c->alloc(&c->rng, (size_t)B * 8));
    RCHECK(c->alloc(&c->acc_count, (size_t)B * 8));
    RCHECK(c->alloc(&c->accmask, (size_t)B * L * ((N + 31) / 32) * 4));
    RCHECK(c->alloc(&c->fields, (size_t)B * L * N));
    RCHECK(c->alloc(&c->G64, (size_t)B * nn * 8));
    if (precision == DQMC_MIXED) {
        RCHECK(c->alloc(&c->G32, (size_t)B * nn * 4)); RCHECK(c->alloc(&c->T32, (size_t)B * nn * 4));
        RCHECK(c->alloc(&c->expK32, nn * 4)); RCHECK(c->alloc(&c->invExpK32, nn * 4));
    }
    RCHECK(c->alloc(&c->stkA, (size_t)B * nn * 8)); RCHECK(c->alloc(&c->stkB, (size_t)B * nn * 8));
    RCHECK(c->alloc(&c->W1, (size_t)B * nn * 8)); RCHECK(c->alloc(&c->W2, (size_t)B * nn * 8));
    RCHECK(c->alloc(&c->W3, (size_t)B * nn * 8));
    RCHECK(c->alloc(&c->qr, (size_t)B * c->sqr() * 8));
    RCHECK(c->alloc(&c->perm, (size_t)B * N * 4));
    RCHECK(c->alloc(&c->lu_perm, (size_t)B * N * 4));
    RCHECK(c->alloc(&c->qrV, (size_t)B * N * 32 * 8)); RCHECK(c->alloc(&c->qrT, (size_t)B * 32 * 32 * 8));
    RCHECK(c->alloc(&c->qrV2, (size_t)B * N * 32 * 8)); RCHECK(c->alloc(&c->qrT2, (size_t)B * 32 * 32 * 8));
    RCHECK(c->alloc(&c->lu_perm2, (size_t)B * N * 4));
    RCHECK(c->alloc(&c->qrWt, (size_t)2 * B * c->wdyn() * 32 * 8));
    for (DStack* s : {&c->left, &c->right}) {
        RCHECK(c->alloc(&s->U, (size_t)B * S * nn * 8));
        RCHECK(c->alloc(&s->V, (size_t)B * S * nn * 8));
        RCHECK(c->alloc(&s->D, (size_t)B * S * N * 8));
    }
    if (c->eqtime || c->dynamic) RCHECK(c->alloc(&c->vecgftt, (size_t)B * L * nn * 8));
    if (c->dynamic) {
        RCHECK(c->alloc(&c->vecgft0, (size_t)B * L * nn * 8)); RCHECK(c->alloc(&c->vecgf0t, (size_t)B * L * nn * 8));
        RCHECK(c->alloc(&c->gt0A, (size_t)B * nn * 8)); RCHECK(c->alloc(&c->gt0B, (size_t)B * nn * 8));
        RCHECK(c->alloc(&c->g0tA, (size_t)B * nn * 8)); RCHECK(c->alloc(&c->g0tB, (size_t)B * nn * 8));
    }
    RCHECK(c->alloc(&c->wrap_eq, (size_t)B * 8)); RCHECK(c->alloc(&c->wrap_dyn, (size_t)B * 8));
    RCHECK(c->alloc(&c->maxdG, (size_t)B * 8));
    RCHECK(c->alloc(&c->health, (size_t)B * 4)); RCHECK(c->alloc(&c->wd_bound, 8));
    RCHECK(c->alloc(&c->cmp_part, (size_t)B * CMP_PART_SLOTS * 3 * 8));
    RCHECK(c->alloc(&c->colnorm, (size_t)B * N * 8));

    // constant tables
    std::vector<double> ek, iek;
    lattice_expk(c->nl, c->dt * c->nnt, ek);    // exp(-dt K) at mu = 0
    lattice_expk(c->nl, -c->dt * c->nnt, iek);  // exp(+dt K) at mu = 0
    HCHECK(c, hipMemcpy(c->expK, ek.data(), nn * 8, hipMemcpyHostToDevice));
    HCHECK(c, hipMemcpy(c->invExpK, iek.data(), nn * 8, hipMemcpyHostToDevice));
    {
        std::vector<double> id((size_t)nn, 0.0);
        for (int i = 0; i < N; ++i) id[(size_t)i * N + i] = 1.0;
        HCHECK(c, hipMemcpy(c->ident, id.data(), nn * 8, hipMemcpyHostToDevice));
    }
    if (precision == DQMC_MIXED) {
        std::vector<float> f(nn);
        for (long long i = 0; i < nn; ++i) f[i] = (float)ek[i];
        HCHECK(c, hipMemcpy(c->expK32, f.data(), nn * 4, hipMemcpyHostToDevice));
        for (long long i = 0; i < nn; ++i) f[i] = (float)iek[i];
        HCHECK(c, hipMemcpy(c->invExpK32, f.data(), nn * 4, hipMemcpyHostToDevice));
    }
    {
        // 16 x 16 padded up to 16 sites per side, 32 x 32 above (the FP32 copies serve nl <= 16 only)
        const int P = c->nl > 16 ? 32 : 16;
        std::vector<double> e1(1024), f1(1024);
        ring_expk_padded(c->nl, c->dt * c->nnt, P, e1.data());
        ring_expk_padded(c->nl, -c->dt * c->nnt, P, f1.data());
        std::vector<float> e1f(256), f1f(256);
        for (int i = 0; i < 256; ++i) { e1f[i] = (float)e1[i]; f1f[i] = (float)f1[i]; }
        HCHECK(c, hipMemcpy(c->E1, e1.data(), 1024 * 8, hipMemcpyHostToDevice));
        HCHECK(c, hipMemcpy(c->F1, f1.data(), 1024 * 8, hipMemcpyHostToDevice));
        HCHECK(c, hipMemcpy(c->E1f, e1f.data(), 256 * 4, hipMemcpyHostToDevice));
        HCHECK(c, hipMemcpy(c->F1f, f1f.data(), 256 * 4, hipMemcpyHostToDevice));
    }
    std::vector<uint64_t> pw(2 * N);
    uint64_t x = 1;
    for (int k = 0; k < 2 * N; ++k) { x = (x * 16807ULL) % kMinstdM; pw[k] = x; }
    HCHECK(c, hipMemcpy(c->powtab, pw.data(), pw.size() * 8, hipMemcpyHostToDevice));
    std::vector<unsigned long long> zero(B, 0ULL);
    HCHECK(c, hipMemcpy(c->acc_count, zero.data(), (size_t)B * 8, hipMemcpyHostToDevice));
    std::vector<int8_t> ones((size_t)B * L * N, 1);
    HCHECK(c, hipMemcpy(c->fields, ones.data(), ones.size(), hipMemcpyHostToDevice));
    std::vector<uint64_t> r1(B, 1ULL);
    HCHECK(c, hipMemcpy(c->rng, r1.data(), (size_t)B * 8, hipMemcpyHostToDevice));
    const double off = 0.0;
    HCHECK(c, hipMemcpy(c->wd_bound, &off, 8, hipMemcpyHostToDevice));
    std::vector<uint32_t> h0(B, 0u);
    HCHECK(c, hipMemcpy(c->health, h0.data(), (size_t)B * 4, hipMemcpyHostToDevice));
    return 0;
}

void dqmc_destroy(dqmc_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
    for (void* p : c->allocs) (void)hipFree(p);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : c->kt_ev) (void)hipEventDestroy(e);
    if (c->stream2) (void)hipStreamSynchronize(c->stream2);
    if (c->la_fork) (void)hipEventDestroy(c->la_fork);
    if (c->la_join) (void)hipEventDestroy(c->la_join);
    if (c->ms_fork) (void)hipEventDestroy(c->ms_fork);
    for (int g = 1; g < MSPLIT_MAX; ++g)
        if (c->ms_join[g]) (void)hipEventDestroy(c->ms_join[g]);
    for (int g = 2; g < MSPLIT_MAX; ++g)
        if (c->ms_stream[g]) (void)hipStreamDestroy(c->ms_stream[g]);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* dqmc_last_error(dqmc_ctx* c) { return c ? c->err.c_str() : dqmc::free_error_text(); }

static int check_chain(dqmc_ctx* c, int chain)
{
    if (!c) return DQMC_EINVAL;
    if (chain < 0 || chain >= c->B) return c->fail(DQMC_EINVAL, "chain %d out of range [0,%d)", chain, c->B);
    return 0;
}

int dqmc_set_collect(dqmc_ctx* c, int eq, int dyn)
{
    if (!c) return DQMC_EINVAL;
    if (c->initialized) return c->fail(DQMC_ESTATE, "set_collect after init");
    const long long nn = c->NN();
    const size_t B = c->B, L = c->L;
    c->eqtime = eq != 0;
    c->dynamic = dyn != 0;
    if ((c->eqtime || c->dynamic) && !c->vecgftt) RCHECK(c->alloc(&c->vecgftt, B * L * nn * 8));
    if (c->dynamic && !c->vecgft0) {
        RCHECK(c->alloc(&c->vecgft0, B * L * nn * 8)); RCHECK(c->alloc(&c->vecgf0t, B * L * nn * 8));
        RCHECK(c->alloc(&c->gt0A, B * nn * 8)); RCHECK(c->alloc(&c->gt0B, B * nn * 8));
        RCHECK(c->alloc(&c->g0tA, B * nn * 8)); RCHECK(c->alloc(&c->g0tB, B * nn * 8));
    }
    return 0;
}

int dqmc_set_chain_mu(dqmc_ctx* c, int chain, double mu)
{
    RCHECK(check_chain(c, chain));
    if (c->initialized) return c->fail(DQMC_ESTATE, "set_chain_mu after init");
    c->chain_mu[chain] = mu;
    return 0;
}

int dqmc_set_seed(dqmc_ctx* c, int chain, uint64_t seed)
{
    RCHECK(check_chain(c, chain));
    uint64_t v = seed % kMinstdM;
    c->rng_host[chain] = v ? v : 1ULL;   // libstdc++ linear_congruential_engine::seed
    HCHECK(c, hipMemcpy(c->rng + chain, &c->rng_host[chain], 8, hipMemcpyHostToDevice));
    return 0;
}

int dqmc_get_rng_state(dqmc_ctx* c, int chain, uint64_t* state)
{
    RCHECK(check_chain(c, chain));
    if (!state) return DQMC_EINVAL;
    HCHECK(c, hipMemcpy(state, c->rng + chain, 8, hipMemcpyDeviceToHost));
    c->rng_host[chain] = *state;
    return 0;
}

int dqmc_fields_random(dqmc_ctx* c, int chain)
{
    RCHECK(check_chain(c, chain));
    uint64_t x = 0;
    HCHECK(c, hipMemcpy(&x, c->rng + chain, 8, hipMemcpyDeviceToHost));
    std::vector<int8_t> s((size_t)c->L * c->N);
    for (auto& v : s) v = host_canonical(x) < 0.5 ? 1 : -1;
    HCHECK(c, hipMemcpy(c->fields + (size_t)chain * c->L * c->N, s.data(), s.size(), hipMemcpyHostToDevice));
    HCHECK(c, hipMemcpy(c->rng + chain, &x, 8, hipMemcpyHostToDevice));
    c->rng_host[chain] = x;
    return 0;
}

int dqmc_set_fields(dqmc_ctx* c, int chain, const int8_t* s)
{
    RCHECK(check_chain(c, chain));
    if (!s) return DQMC_EINVAL;
    for (long long k = 0; k < (long long)c->L * c->N; ++k)
        if (s[k] != 1 && s[k] != -1) return c->fail(DQMC_EINVAL, "field value %d at %lld is not +1/-1", s[k], k);
    HCHECK(c, hipMemcpy(c->fields + (size_t)chain * c->L * c->N, s, (size_t)c->L * c->N, hipMemcpyHostToDevice));
    return 0;
}

int dqmc_get_fields(dqmc_ctx* c, int chain, int8_t* s)
{
    RCHECK(check_chain(c, chain));
    if (!s) return DQMC_EINVAL;
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(s, c->fields + (size_t)chain * c->L * c->N, (size_t)c->L * c->N, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_init(dqmc_ctx* c)
{
    if (!c) return DQMC_EINVAL;
    const int B = c->B, N = c->N, L = c->L;
    const long long nn = c->NN();
    // per-chain scale tables: B = diag(e^{dt mu + alpha s}) exp(dt nnt A), B^{-1} = exp(-dt nnt A) diag(e^{-dt mu - alpha s})
    std::vector<double> tab(2 * B), itab(2 * B), del(2 * B);
    for (int b = 0; b < B; ++b) {
        const double m = c->dt * c->chain_mu[b];
        tab[2 * b] = std::exp(m + c->alpha); tab[2 * b + 1] = std::exp(m - c->alpha);
        itab[2 * b] = std::exp(-m - c->alpha); itab[2 * b + 1] = std::exp(-m + c->alpha);
        del[2 * b] = std::exp(-2 * c->alpha * (+1)) - 1;   // src/hubbard.cpp:127
        del[2 * b + 1] = std::exp(-2 * c->alpha * (-1)) - 1;
    }
    HCHECK(c, hipMemcpy(c->tab, tab.data(), 16 * (size_t)B, hipMemcpyHostToDevice));
    HCHECK(c, hipMemcpy(c->itab, itab.data(), 16 * (size_t)B, hipMemcpyHostToDevice));
    HCHECK(c, hipMemcpy(c->delta, del.data(), 16 * (size_t)B, hipMemcpyHostToDevice));
    HCHECK(c, launch_fill_d(c->wrap_eq, B, 0.0, c->stream));
    HCHECK(c, launch_fill_d(c->wrap_dyn, B, 0.0, c->stream));
    HCHECK(c, launch_fill_d(c->maxdG, B, 0.0, c->stream));
    if (!c->initialized) {
        RCHECK(autotune(c));
        const char* gt = std::getenv("DQMC_GEMM_TILE");   // override of the autotuner's choice
        if (gt && gt[0] >= '0' && gt[0] < '0' + TILE_COUNT && gt[1] == 0) c->tile64 = c->tile32 = gt[0] - '0';
    }
    RCHECK(health_reset(c));
    RCHECK(build_from_fields(c));
    HCHECK(c, hipStreamSynchronize(c->stream));
    (void)N; (void)L; (void)nn;
    c->initialized = true;
    return health_check(c, "dqmc_init");
}

int dqmc_apply_bop(dqmc_ctx* c, int chain, int op, int t, double* inout)
{
    RCHECK(check_chain(c, chain));
    if (!inout || op < 0 || op > 4 || t < 0 || t >= c->L) return c->fail(DQMC_EINVAL, "bad apply_bop arguments");
    if (!c->initialized) return c->fail(DQMC_ESTATE, "apply_bop before init");
    const long long nn = c->NN();
    // every chain computes on W2 (only `chain`'s result is read back)
    HCHECK(c, hipMemcpy(c->W2 + chain * nn, inout, nn * 8, hipMemcpyHostToDevice));
    switch (op) {
        case 0: RCHECK(left_expk(c, c->W2, c->W3, c->tab, t, false)); break;    // D e^{-dtK} G
        case 1: RCHECK(right_expk(c, c->W2, c->W3, c->tab, t, true)); break;    // G D e^{-dtK}
        case 2: RCHECK(left_expk(c, c->W2, c->W3, c->itab, t, true)); break;    // e^{dtK} D^{-1} G
        case 3: RCHECK(right_expk(c, c->W2, c->W3, c->itab, t, false)); break;  // G e^{dtK} D^{-1}
        default: RCHECK(left_expk(c, c->W2, c->W3, c->tab, t, true)); break;    // e^{-dtK} D G
    }
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(inout, c->W3 + chain * nn, nn * 8, hipMemcpyDeviceToHost));
    return 0;
}

// ---- single-site operations of Model::HubbardAttractiveU (include/hubbard.h:71-73) ----------
static int site_args(dqmc_ctx* c, int chain, int t, int i, const char* what)
{
    RCHECK(check_chain(c, chain));
    if (t < 0 || t >= c->L || i < 0 || i >= c->N) return c->fail(DQMC_EINVAL, "%s: (t, i) out of range", what);
    if (!c->initialized) return c->fail(DQMC_ESTATE, "%s before dqmc_init", what);
    if (c->prec != DQMC_FP64) return c->fail(DQMC_EINVAL, "%s needs an FP64 context", what);
    return 0;
}
static int site_delta(dqmc_ctx* c, int chain, int t, int i, double* delta)
{
    int8_t s = 0;
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(&s, c->fields + ((size_t)chain * c->L + t) * c->N + i, 1, hipMemcpyDeviceToHost));
    *delta = std::exp(-2 * c->alpha * s) - 1;   // src/hubbard.cpp:127 (the table of dqmc_init)
    return 0;
}

int dqmc_acceptance_ratio(dqmc_ctx* c, int chain, int t, int i, double* ratio)
{
#pragma clang fp contract(off)
    RCHECK(site_args(c, chain, t, i, "dqmc_acceptance_ratio"));
    if (!ratio) return DQMC_EINVAL;
    double delta = 0.0, gii = 0.0;
    RCHECK(site_delta(c, chain, t, i, &delta));
    HCHECK(c, hipMemcpy(&gii, c->G64 + (size_t)chain * c->NN() + (size_t)i * c->N + i, 8, hipMemcpyDeviceToHost));
    const double x = 1 + (1 - gii) * delta;   // src/hubbard.cpp:123-130
    *ratio = x * x / (1 + delta);
    return 0;
}

int dqmc_update_green_function(dqmc_ctx* c, int chain, int t, int i)
{
    RCHECK(site_args(c, chain, t, i, "dqmc_update_green_function"));
    double delta = 0.0;
    RCHECK(site_delta(c, chain, t, i, &delta));
    HCHECK(c, launch_site_update(c->G64 + (size_t)chain * c->NN(), c->N, i, delta, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dqmc_flip_field(dqmc_ctx* c, int chain, int t, int i)
{
    RCHECK(check_chain(c, chain));
    if (t < 0 || t >= c->L || i < 0 || i >= c->N) return c->fail(DQMC_EINVAL, "dqmc_flip_field: (t, i) out of range");
    int8_t s = 0;
    int8_t* p = c->fields + ((size_t)chain * c->L + t) * c->N + i;
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(&s, p, 1, hipMemcpyDeviceToHost));
    s = (int8_t)-s;
    HCHECK(c, hipMemcpy(p, &s, 1, hipMemcpyHostToDevice));
    return 0;
}

int dqmc_set_thermalization(dqmc_ctx* c, int th)
{
    if (!c) return DQMC_EINVAL;
    c->thermalization = th != 0;
    return 0;
}

int dqmc_sweep_up(dqmc_ctx* c)
{
    if (!c) return DQMC_EINVAL;
    if (!c->initialized || c->current_t != 0 || c->left.len != 0 || c->right.len != c->S)
        return c->fail(DQMC_ESTATE, "sweep_up needs current_t == 0, empty left and full right stacks");
    RCHECK(health_reset(c));
    RCHECK(run_sweep(c, do_sweep_up, 0));
    c->current_t = c->L;
    c->updating_sweeps += 1;
    return health_check(c, "dqmc_sweep_up");
}

int dqmc_sweep_down(dqmc_ctx* c)
{
    if (!c) return DQMC_EINVAL;
    if (!c->initialized || c->current_t != c->L || c->right.len != 0 || c->left.len != c->S)
        return c->fail(DQMC_ESTATE, "sweep_down needs current_t == nt, empty right and full left stacks");
    RCHECK(health_reset(c));
    RCHECK(run_sweep(c, do_sweep_down, 1));
    c->current_t = 0;
    c->updating_sweeps += 1;
    return health_check(c, "dqmc_sweep_down");
}

int dqmc_sweep_dynamic(dqmc_ctx* c)
{
    if (!c) return DQMC_EINVAL;
    if (!c->dynamic) return 0;   // Core::sweep_for_dynamic_green_functions is a no-op then
    if (!c->initialized || c->current_t != 0 || c->left.len != 0 || c->right.len != c->S)
        return c->fail(DQMC_ESTATE, "sweep_dynamic needs current_t == 0, empty left and full right stacks");
    RCHECK(health_reset(c));
    RCHECK(run_sweep(c, do_sweep_dynamic, 2));
    c->current_t = c->L;
    return health_check(c, "dqmc_sweep_dynamic");
}

int dqmc_get_gf(dqmc_ctx* c, int chain, int kind, int t, double* out)
{
    RCHECK(check_chain(c, chain));
    if (!out) return DQMC_EINVAL;
    const long long nn = c->NN();
    HCHECK(c, hipStreamSynchronize(c->stream));
    if (kind == DQMC_GF_CURRENT) {
        if (c->prec == DQMC_MIXED) {
            std::vector<float> f(nn);
            HCHECK(c, hipMemcpy(f.data(), c->G32 + chain * nn, nn * 4, hipMemcpyDeviceToHost));
            for (long long i = 0; i < nn; ++i) out[i] = f[i];
        } else {
            HCHECK(c, hipMemcpy(out, c->G64 + chain * nn, nn * 8, hipMemcpyDeviceToHost));
        }
        return 0;
    }
    if (t < 0 || t >= c->L) return c->fail(DQMC_EINVAL, "t out of range");
    const double* src = kind == DQMC_GF_TT ? c->vecgftt : kind == DQMC_GF_T0 ? c->vecgft0 : kind == DQMC_GF_0T ? c->vecgf0t : nullptr;
    if (!src) return c->fail(DQMC_EINVAL, "Green's function kind %d not collected", kind);
    HCHECK(c, hipMemcpy(out, src + ((long long)chain * c->L + t) * nn, nn * 8, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_set_gf(dqmc_ctx* c, int chain, const double* g)
{
    RCHECK(check_chain(c, chain));
    if (!g) return DQMC_EINVAL;
    const long long nn = c->NN();
    HCHECK(c, hipMemcpy(c->G64 + chain * nn, g, nn * 8, hipMemcpyHostToDevice));
    if (c->prec == DQMC_MIXED) {
        std::vector<float> f(nn);
        for (long long i = 0; i < nn; ++i) f[i] = (float)g[i];
        HCHECK(c, hipMemcpy(c->G32 + chain * nn, f.data(), nn * 4, hipMemcpyHostToDevice));
    }
    return 0;
}

int dqmc_get_stats(dqmc_ctx* c, int chain, double* wrap_eq, double* wrap_dyn, double* max_abs_dG, long long* accepts,
                   long long* visits)
{
    RCHECK(check_chain(c, chain));
    HCHECK(c, hipStreamSynchronize(c->stream));
    double v;
    if (wrap_eq) { HCHECK(c, hipMemcpy(&v, c->wrap_eq + chain, 8, hipMemcpyDeviceToHost)); *wrap_eq = v; }
    if (wrap_dyn) { HCHECK(c, hipMemcpy(&v, c->wrap_dyn + chain, 8, hipMemcpyDeviceToHost)); *wrap_dyn = v; }
    if (max_abs_dG) { HCHECK(c, hipMemcpy(&v, c->maxdG + chain, 8, hipMemcpyDeviceToHost)); *max_abs_dG = v; }
    unsigned long long a = 0;
    HCHECK(c, hipMemcpy(&a, c->acc_count + chain, 8, hipMemcpyDeviceToHost));
    if (accepts) *accepts = (long long)a;
    if (visits) *visits = c->updating_sweeps * (long long)c->L * c->N;
    return 0;
}

int dqmc_get_accept_mask(dqmc_ctx* c, int chain, uint32_t* out)
{
    RCHECK(check_chain(c, chain));
    if (!out) return DQMC_EINVAL;
    const size_t n = (size_t)c->L * ((c->N + 31) / 32);
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(out, c->accmask + chain * n, n * 4, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_synchronize(dqmc_ctx* c)
{
    if (!c) return DQMC_EINVAL;
    HCHECK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int dqmc_event_record(dqmc_ctx* c, int slot)
{
    if (!c || slot < 0 || slot >= 8) return DQMC_EINVAL;
    HCHECK(c, hipEventRecord(c->ev[slot], c->stream));
    return 0;
}

int dqmc_event_elapsed(dqmc_ctx* c, int a, int b, float* ms)
{
    if (!c || a < 0 || a >= 8 || b < 0 || b >= 8 || !ms) return DQMC_EINVAL;
    HCHECK(c, hipEventSynchronize(c->ev[b]));
    HCHECK(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return 0;
}

int dqmc_kernel_timer_begin(dqmc_ctx* c, int kernel, int max_launches)
{
    if (!c || kernel < 0 || kernel > 4 || max_launches < 1) return DQMC_EINVAL;
    while ((int)c->kt_ev.size() < 2 * max_launches) {
        hipEvent_t e;
        HCHECK(c, hipEventCreate(&e));
        c->kt_ev.push_back(e);
    }
    c->kt_kernel = kernel;
    c->kt_used = 0;
    return 0;
}

int dqmc_kernel_timer_end(dqmc_ctx* c, double* total_ms, int* launches)
{
    if (!c) return DQMC_EINVAL;
    HCHECK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (int i = 0; i + 1 < c->kt_used; i += 2) {
        float ms = 0.f;
        HCHECK(c, hipEventElapsedTime(&ms, c->kt_ev[i], c->kt_ev[i + 1]));
        tot += ms;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = c->kt_used / 2;
    c->kt_kernel = -1;
    c->kt_used = 0;
    return 0;
}

int dqmc_use_graphs(dqmc_ctx* c, int on)
{
    if (!c) return DQMC_EINVAL;
    c->use_graphs = on != 0;
    return 0;
}

int dqmc_get_tuning(dqmc_ctx* c, int* tile_fp64, int* tile_fp32, float* ms)
{
    if (!c) return DQMC_EINVAL;
    if (!c->initialized) return c->fail(DQMC_ESTATE, "dqmc_get_tuning: the autotuner runs in dqmc_init");
    if (tile_fp64) *tile_fp64 = c->tile64;
    if (tile_fp32) *tile_fp32 = c->tile32;
    if (ms)
        for (int p = 0; p < 2; ++p)
            for (int t = 0; t < TILE_COUNT; ++t) ms[p * TILE_COUNT + t] = c->tune_ms[p][t];
    return 0;
}

int dqmc_set_tuning(dqmc_ctx* c, int tile_fp64, int tile_fp32)
{
    if (!c) return DQMC_EINVAL;
    if (tile_fp64 < 0 || tile_fp64 >= TILE_COUNT || tile_fp32 < 0 || tile_fp32 >= TILE_COUNT)
        return c->fail(DQMC_EINVAL, "dqmc_set_tuning: tile index out of range");
    if (!c->initialized) return c->fail(DQMC_ESTATE, "dqmc_set_tuning: call after dqmc_init");
    if (!c->graphs.empty()) return c->fail(DQMC_ESTATE, "dqmc_set_tuning: sweeps were already captured");
    c->tile64 = tile_fp64;
    c->tile32 = tile_fp32;
    return 0;
}

int dqmc_forward_fixed(dqmc_ctx* c, double* logdet, double* gdiag)
{
    if (!c) return DQMC_EINVAL;
    if (!c->initialized) return c->fail(DQMC_ESTATE, "dqmc_forward_fixed before dqmc_init");
    if (c->prec != DQMC_FP64) return c->fail(DQMC_EINVAL, "dqmc_forward_fixed needs an FP64 context");
    const int B = c->B, N = c->N, L = c->L;
    const long long nn = c->NN();
    RCHECK(ensure_adjoint_buffers(c));
    SavedStats st;
    RCHECK(save_stats(c, st));
    RCHECK(health_reset(c));
    RCHECK(build_from_fields(c));
    HCHECK(c, launch_logdet_00bb(c->qr, c->sqr(), c->weq(), sD(c, c->right, c->right.len - 1), (long long)c->S * N, N,
                                 c->adj_out, B, c->stream));
    RCHECK(do_sweep_dynamic(c));
    HCHECK(c, launch_gf_diag(c->vecgftt, (long long)L * nn, L, N, c->adj_gdiag, B, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    if (logdet) HCHECK(c, hipMemcpy(logdet, c->adj_out, (size_t)B * 8, hipMemcpyDeviceToHost));
    if (gdiag) HCHECK(c, hipMemcpy(gdiag, c->adj_gdiag, (size_t)B * L * N * 8, hipMemcpyDeviceToHost));
    RCHECK(build_from_fields(c));
    RCHECK(restore_stats(c, st));
    return health_check(c, "dqmc_forward_fixed");
}

int dqmc_adjoint(dqmc_ctx* c, const double* gbar_logdet, const double* gbar_gdiag, double* grad)
{
    if (!c || !grad) return DQMC_EINVAL;
    if (!c->initialized) return c->fail(DQMC_ESTATE, "dqmc_adjoint before dqmc_init");
    if (c->prec != DQMC_FP64) return c->fail(DQMC_EINVAL, "dqmc_adjoint needs an FP64 context");
    const int B = c->B, N = c->N, L = c->L;
    const long long nn = c->NN();
    RCHECK(ensure_adjoint_buffers(c));
    SavedStats st;
    RCHECK(save_stats(c, st));
    std::vector<double> zl;
    if (!gbar_logdet) { zl.assign(B, 0.0); gbar_logdet = zl.data(); }
    HCHECK(c, hipMemcpy(c->adj_glogdet, gbar_logdet, (size_t)B * 8, hipMemcpyHostToDevice));
    if (gbar_gdiag) HCHECK(c, hipMemcpy(c->adj_gdiag, gbar_gdiag, (size_t)B * L * N * 8, hipMemcpyHostToDevice));
    else HCHECK(c, launch_fill_d(c->adj_gdiag, (long long)B * L * N, 0.0, c->stream));
    HCHECK(c, launch_fill_d(c->adj_grad, (long long)B * 3, 0.0, c->stream));
    RCHECK(health_reset(c));
    HCHECK(c, hipMemcpyAsync(c->fields_save, c->fields, (size_t)B * L * N, hipMemcpyDeviceToDevice, c->stream));
    // d alpha / d u for cosh(alpha) = e^{dt u / 2} (src/hubbard.cpp:35)
    const double dalpha = c->dt * std::exp(0.5 * c->dt * c->u) / (2.0 * std::sinh(c->alpha));
    for (int tau = 0; tau < L; ++tau) {
        HCHECK(c, launch_roll_fields(c->fields_save, c->fields, L, N, tau, B, c->stream));
        RCHECK(build_from_fields(c));
        RCHECK(do_sweep_dynamic(c));
        AdjointArgs a;
        a.gf0t = c->vecgf0t; a.gft0 = c->vecgft0; a.gftt = c->vecgftt; a.sv = (long long)L * nn;
        a.gbar_diag = c->adj_gdiag; a.gbar_logdet = c->adj_glogdet; a.fields = c->fields; a.nb = c->nb;
        a.partial = c->adj_partial; a.grad = c->adj_grad;
        a.N = N; a.L = L; a.tau = tau;
        a.alpha = c->alpha; a.dt = c->dt; a.dalpha = dalpha;
        HCHECK(c, launch_adjoint_contract(a, B, c->stream));
    }
    HCHECK(c, hipMemcpyAsync(c->fields, c->fields_save, (size_t)B * L * N, hipMemcpyDeviceToDevice, c->stream));
    RCHECK(build_from_fields(c));
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(grad, c->adj_grad, (size_t)B * 3 * 8, hipMemcpyDeviceToHost));
    RCHECK(restore_stats(c, st));
    return health_check(c, "dqmc_adjoint");
}

int dqmc_set_watchdog(dqmc_ctx* c, double max_dG)
{
    if (!c) return DQMC_EINVAL;
    if (!(max_dG >= 0.0)) return c->fail(DQMC_EINVAL, "dqmc_set_watchdog: bound must be >= 0 (0 = off)");
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(c->wd_bound, &max_dG, 8, hipMemcpyHostToDevice));
    return 0;
}

int dqmc_get_health(dqmc_ctx* c, int chain, uint32_t* flags)
{
    RCHECK(check_chain(c, chain));
    if (!flags) return DQMC_EINVAL;
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(flags, c->health + chain, 4, hipMemcpyDeviceToHost));
    return 0;
}

int dqmc_measure_setup(dqmc_ctx* c, unsigned obs_mask, int nbins, int nk, const double* kxy)
{
    if (!c) return DQMC_EINVAL;
    if (obs_mask == 0 || (obs_mask & ~63u) || nbins < 1 || nk < 0 || (nk > 0 && !kxy))
        return c->fail(DQMC_EINVAL, "dqmc_measure_setup: bad observable mask, bin count or momentum list");
    if ((obs_mask & DQMC_OBS_GREEN_KT) && nk < 1)
        return c->fail(DQMC_EINVAL, "dqmc_measure_setup: DynamicGreenFunctions needs a momentum list");
    if ((obs_mask & 7u) && !c->eqtime)
        return c->fail(DQMC_ESTATE, "dqmc_measure_setup: equal-time observables need dqmc_set_collect(eqtime)");
    if ((obs_mask & 56u) && !c->dynamic)
        return c->fail(DQMC_ESTATE, "dqmc_measure_setup: dynamic observables need dqmc_set_collect(dynamic)");
    if (c->meas_cache) return c->fail(DQMC_ESTATE, "dqmc_measure_setup: already set up");
    const int B = c->B, L = c->L, N = c->N;
    c->obs_mask = obs_mask; c->nbins = nbins; c->nk = nk;
    c->cstride = 3 + nk * L + 2 * L;
    RCHECK(c->alloc(&c->meas_cache, (size_t)B * c->cstride * 8));
    RCHECK(c->alloc(&c->meas_data, (size_t)B * nbins * c->cstride * 8));
    RCHECK(c->alloc(&c->meas_part, (size_t)B * L * 3 * 8));
    RCHECK(c->alloc(&c->costab, (size_t)std::max(1, nk) * N * 8));
    // Re exp(-i k . r_d) for the displacement of site index d (src/square_lattice.cpp:120-135)
    std::vector<double> ct((size_t)std::max(1, nk) * N, 0.0);
    for (int k = 0; k < nk; ++k)
        for (int d = 0; d < N; ++d) ct[(size_t)k * N + d] = std::cos((d % c->nl) * kxy[2 * k] + (d / c->nl) * kxy[2 * k + 1]);
    HCHECK(c, hipMemcpy(c->costab, ct.data(), ct.size() * 8, hipMemcpyHostToDevice));
    HCHECK(c, launch_fill_d(c->meas_cache, (long long)B * c->cstride, 0.0, c->stream));
    HCHECK(c, launch_fill_d(c->meas_data, (long long)B * nbins * c->cstride, 0.0, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    c->eq_count = c->dyn_count = 0;
    return 0;
}

int dqmc_measure_equaltime(dqmc_ctx* c)
{
    if (!c) return DQMC_EINVAL;
    if (!c->meas_cache || !(c->obs_mask & 7u)) return c->fail(DQMC_ESTATE, "dqmc_measure_equaltime: no equal-time observables set up");
    if (!c->initialized) return c->fail(DQMC_ESTATE, "dqmc_measure_equaltime before dqmc_init");
    HCHECK(c, launch_measure_equaltime(c->vecgftt, (long long)c->L * c->NN(), c->L, c->N, c->meas_part, c->meas_cache,
                                       c->cstride, c->B, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    c->eq_count += 1;
    return 0;
}

int dqmc_measure_dynamic(dqmc_ctx* c)
{
    if (!c) return DQMC_EINVAL;
    if (!c->meas_cache || !(c->obs_mask & 56u)) return c->fail(DQMC_ESTATE, "dqmc_measure_dynamic: no dynamic observables set up");
    if (!c->initialized) return c->fail(DQMC_ESTATE, "dqmc_measure_dynamic before dqmc_init");
    const int flags = ((c->obs_mask & DQMC_OBS_GREEN_KT) ? 1 : 0) | ((c->obs_mask & DQMC_OBS_LDOS) ? 2 : 0) |
                      ((c->obs_mask & DQMC_OBS_SPIN) ? 4 : 0);
    HCHECK(c, launch_measure_dynamic(c->vecgft0, c->vecgf0t, (long long)c->L * c->NN(), c->L, c->nl, c->costab, c->nk,
                                     c->meas_cache, c->cstride, flags, c->B, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    c->dyn_count += 1;
    return 0;
}

int dqmc_measure_push_bin(dqmc_ctx* c, int bin)
{
    if (!c) return DQMC_EINVAL;
    if (!c->meas_cache) return c->fail(DQMC_ESTATE, "dqmc_measure_push_bin: no observables set up");
    if (bin < 0 || bin >= c->nbins)
        return c->fail(DQMC_EINVAL, "Observable::Observable<_scalar>::push_cache_to_data(): index of bin out of range.");
    if (((c->obs_mask & 7u) && c->eq_count == 0) || ((c->obs_mask & 56u) && c->dyn_count == 0))
        return c->fail(DQMC_ESTATE, "Observable::Observable<_scalar>::normalize_cache(): cache divided by zero.");
    HCHECK(c, launch_measure_push(c->meas_cache, c->cstride, c->meas_data, (long long)c->nbins * c->cstride, bin, 3,
                                  c->eq_count, c->dyn_count, c->B, c->stream));
    HCHECK(c, hipStreamSynchronize(c->stream));
    c->eq_count = c->dyn_count = 0;
    return 0;
}

int dqmc_get_bins(dqmc_ctx* c, int chain, unsigned obs, double* out)
{
    RCHECK(check_chain(c, chain));
    if (!out || !c->meas_cache) return c->fail(DQMC_EINVAL, "dqmc_get_bins: no output or no observables set up");
    if (!(obs & c->obs_mask) || (obs & (obs - 1))) return c->fail(DQMC_EINVAL, "dqmc_get_bins: observable %u not measured", obs);
    const int L = c->L;
    std::vector<double> d((size_t)c->nbins * c->cstride);
    HCHECK(c, hipStreamSynchronize(c->stream));
    HCHECK(c, hipMemcpy(d.data(), c->meas_data + (size_t)chain * c->nbins * c->cstride, d.size() * 8,
                        hipMemcpyDeviceToHost));
    int off = 0, len = 1;
    switch (obs) {
        case DQMC_OBS_FILLING: off = 0; break;
        case DQMC_OBS_DOUBLE_OCCU: off = 1; break;
        case DQMC_OBS_SWAVE: off = 2; break;
        case DQMC_OBS_GREEN_KT: off = 3; len = c->nk * L; break;
        case DQMC_OBS_LDOS: off = 3 + c->nk * L; len = L; break;
        default: off = 3 + c->nk * L + L; len = L; break;
    }
    for (int b = 0; b < c->nbins; ++b)
        std::memcpy(out + (size_t)b * len, d.data() + (size_t)b * c->cstride + off, (size_t)len * 8);
    return 0;
}

int dqmc_dims(dqmc_ctx* c, int* ns, int* nt, int* nch)
{
    if (!c) return DQMC_EINVAL;
    if (ns) *ns = c->N;
    if (nt) *nt = c->L;
    if (nch) *nch = c->B;
    return 0;
}

}  // extern "C"

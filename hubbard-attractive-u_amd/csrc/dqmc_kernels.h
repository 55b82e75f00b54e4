// dqmc_kernels.h -- launchers of the gfx950 kernels of the DQMC sweep engine.
//
// Layout conventions (device):
//   * N x N matrices are column-major with leading dimension ld, batched over chains with a
//     fixed element stride between consecutive chains.
//   * Matrices handed to the Householder QR are row-major: element (r, c) at buf[r*W + c]
//     with W = N + P (the P extra columns carry right-hand sides that receive Q^T * rhs).
//   * Fields are int8 [chain][t][i].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dqmc {

// Per-chain numeric health flags, OR-ed in by the kernels that detect them.
enum HealthFlag : uint32_t {
    HEALTH_NONFINITE_G = 1u,   // NaN/Inf in the working G or its FP64 recompute at a stabilization
    HEALTH_ZERO_PIVOT = 2u,    // zero or non-finite |R_ii| in a UDV factorization
    HEALTH_RATIO = 4u,         // non-finite acceptance ratio (non-finite G_ii)
    HEALTH_WATCHDOG = 8u       // max |G_recomputed - G_working| above the watchdog bound
};

// Diagonal scaling operand of a GEMM: value for index i of chain b.
enum DiagMode : int {
    DIAG_NONE = 0,
    DIAG_VEC = 1,       // v[b*vstride + i]
    DIAG_INV = 2,       // 1 / v[...]
    DIAG_INVMAX1 = 3,   // 1 / max(1, v[...])
    DIAG_MIN1 = 4,      // min(1, v[...])
    DIAG_FIELD = 5      // f[b*fstride + i] > 0 ? v[b*vstride + 0] : v[b*vstride + 1]
};

struct DiagArg {
    int mode = DIAG_NONE;
    const double* v = nullptr;
    long long vstride = 0;
    const int8_t* f = nullptr;
    long long fstride = 0;
};

template <typename T>
struct GemmArgs {
    int M, Nc, K;
    const T* A; long long sA; int lda; int tA;   // op(A) is M x K
    const T* B; long long sB; int ldb; int tB;   // op(B) is K x Nc
    T* C; long long sC; int ldc; int tC;         // tC: store C^T (element (i,j) at j + i*ldc... see kernel)
    double alpha, beta;                          // C = alpha * diag(r) op(A) diag(s) op(B) diag(c) + beta * C
    DiagArg rowd, innd, cold;
    int batch;
};

// Kronecker-factored hopping exponential.  The periodic nl x nl square lattice has
// A = A1 (x) I + I (x) A1 (A1 the ring adjacency), so exp(c A) = E (x) E with E = exp(c A1).
// Left:  C = diag(r) (E(x)E) diag(s) X diag(cv)   (acts on the columns of X)
// Right: C = diag(r) X diag(s) (E(x)E) diag(cv)   (acts on the rows of X)
// E is stored padded to 16 x 16 (zeros beyond nl); X and C are N x N column-major, C != X.
template <typename T>
struct KronArgs {
    const T* X; long long sX;
    T* C; long long sC;
    const T* E;          // 16 x 16, row-major, padded
    int nl, N;
    DiagArg rowd, innd, cold;
    int batch;
};
// Several factors of a B-chain block in one pass over the 16 x 16 lattice (FP64):
//   C = B(ts[n-1]) .. B(ts[0]) X diag(cv)                   (adjoint = 0, B(t) = diag(d_t) (E(x)E))
//   C = B(ts[n-1])^T .. B(ts[0])^T X diag(cv)               (adjoint = 1, B(t)^T = (E(x)E) diag(d_t))
// with d_t(i) = table[2 b] when the field of slice t at site i is positive, else table[2 b + 1].
// Each lattice vector (column of X) stays in registers over all n factors.
#define KCHAIN_MAX 16
struct KronChainArgs {
    const double* X; long long sX;
    double* C; long long sC;
    const double* E;                 // 16 x 16, row-major, padded
    const int8_t* fields; long long fstride;   // fields[b * fstride + t * 256 + i]
    const double* table;             // 2 values per chain
    int ts[KCHAIN_MAX];
    int n, adjoint;
    DiagArg cold;
    int batch;
    double* norms = nullptr;         // optional: the squared 2-norms of C's columns, 256 per chain
};
hipError_t launch_kron_chain16(const KronChainArgs& a, hipStream_t s);
// Host: the message of the calling thread's last failed call without a context
// (dqmc_last_error(NULL)); free_error records it and returns code.
int free_error(int code, const char* msg);
const char* free_error_text();

// Host: the factor E = exp(x A1) of the ring of nl sites (x = 0: the identity), padded with zeros to a
// P x P row-major block (P * P entries; nl <= P); ring_expk16 is P = 16 and leaves zeros for nl > 16.
void ring_expk_padded(int nl, double x, int P, double* out);
void ring_expk16(int nl, double x, double* out);
// Lattices of 17..32 sites per side (FP64), E padded to 32 x 32: the left product on 32 x 32 tiles of
// 2 x 2 MFMA blocks; the right product applies (E(x)E) to the rows through two batched transposes
// into S1 and S2 (N x N per chain at the chain stride of X): C^T = Left(X^T) with the row and column
// scalings exchanged.
hipError_t launch_kron32_left(const KronArgs<double>& a, hipStream_t s);
hipError_t launch_kron32_right(const KronArgs<double>& a, double* S1, double* S2, hipStream_t s);
// Y_b = X_b^T for N x N column-major matrices, through 32 x 32 LDS tiles.
hipError_t launch_transpose(const double* X, long long sX, double* Y, long long sY, int N, int batch, hipStream_t s);
template <typename T>
hipError_t launch_kron_left(const KronArgs<T>& a, hipStream_t s);
template <typename T>
hipError_t launch_kron_right(const KronArgs<T>& a, hipStream_t s);

// Tile shapes the autotuner chooses from.
// TILE_64x64_PF: 64 x 64 with double-buffered LDS and global loads two K steps ahead.
// TILE_128x128: FP64 128 x 128 (64 x 64 per wave) for products with both output sides >= 128 and
// K >= 64, TILE_64x64_PF otherwise.
enum GemmTile : int { TILE_64x64 = 0, TILE_32x32 = 1, TILE_64x32 = 2, TILE_64x64_PF = 3, TILE_128x128 = 4,
                      TILE_COUNT = 5 };

template <typename T>
hipError_t launch_gemm(const GemmArgs<T>& a, int tile, hipStream_t s);
// The second product of a two-product FP64 GEMM (same shape and transposes as the first, K each),
//   C = diag(a.rowd) op(A) op(B) diag(a.cold) + diag(rowd2) op(A2) op(B2) diag(cold2),
// and the optional row-major copy R (ld ldr) of the first scaled operand diag(a.rowd) op(A).
struct GemmPair {
    const double* A2 = nullptr; long long sA2 = 0; int lda2 = 0;
    const double* B2 = nullptr; long long sB2 = 0; int ldb2 = 0;
    DiagArg rowd2, cold2;
    double* R = nullptr; long long sR = 0; int ldr = 0;
};
// One 128 x 128-tile launch of the two products (op(A) = A^T, op(B) = B, beta = 0, no inner scaling;
// M, Nc >= 128 and K >= 64), the scalings applied to the operands in LDS.
hipError_t launch_gemm_pair(const GemmArgs<double>& a, const GemmPair& p, hipStream_t s);
// C = op(A) diag(innd) op(B) (op(A) = A, op(B) = B^T, 128 x 128 FP64 tile, at most max_tiles tiles per chain)
// stored over the previous C after comparing with it: per tile the signed maximum and maximum absolute
// difference (new - old) and a non-finite flag, into part (chain-major, 3 doubles per tile), which
// launch_compare_fold reduces with np = tiles per chain.
hipError_t launch_gemm_commit(const GemmArgs<double>& a, double* part, int max_tiles, hipStream_t s);

// Metropolis sweep of one slice with delayed rank-k updates.
template <typename T>
struct MetroArgs {
    T* G; long long sG;               // working G, column-major N x N per chain
    int8_t* fields;                   // [chain][L][N]
    int N, L, t;
    uint64_t* rng;                    // [chain] minstd_rand0 state
    const uint64_t* powtab;           // 16807^(k+1) mod (2^31-1), k = 0..2N-1
    const double* delta;              // [chain][2]: e^{-2 alpha s} - 1 for s = +1, -1
    uint32_t* accmask;                // [chain][L][ceil(N/32)]
    unsigned long long* acc_count;    // [chain]
    uint32_t* health;                 // [chain] HEALTH_* flags
    int kmax;
    int chain0 = 0;                   // workgroup w runs chain chain0 + w
    // Optional slice wrap, done by each chain's workgroup after its slice (FP64, N = 256):
    //   left product  lout = diag(r) (E(x)E) diag(s) G with the diagonal of slice lt from ltab as r
    //                 (lmode 1) or s (lmode 2); lmode 0 skips the wrap;
    //   right product G = lout diag(s') (F(x)F) diag(c') with the diagonal of slice lt from rtab as c'
    //                 (rmode 1) or s' (rmode 2), factor rE; rmode 0 leaves lout as the result.
    double* lout = nullptr;
    const double* lE = nullptr;
    const double* ltab = nullptr;
    int lt = 0, lmode = 0;
    const double* rE = nullptr;
    const double* rtab = nullptr;
    int rmode = 0;
};
// threads: workgroup size for N >= 256 (256, 512 or 1024); smaller lattices use 2 x N rounded
// up to a wave.
template <typename T>
hipError_t launch_metropolis(const MetroArgs<T>& a, int batch, int threads, hipStream_t s);
// LDS bytes of one Metropolis workgroup (krows pending-update rows of X and Y).
size_t metropolis_lds_bytes(int N, int krows, size_t esz);
// One-time kernel attribute setup (large dynamic LDS); call before any graph capture.
hipError_t prepare_kernels();

// Householder QR (no pivoting) of the leading N x N block of a row-major N x W buffer,
// applying Q^T to the trailing W-N columns.  On exit the leading block holds R (zeros
// below the diagonal).  One-kernel form (DQMC_QR=v1).
hipError_t launch_qr_v1(double* buf, long long stride, int N, int W, int batch, hipStream_t s);
// Blocked Householder QR of the same row-major N x W system (N <= 1024): per panel of 32 columns
// one factorization launch (one workgroup per chain) and the trailing update of the remaining
// columns.  Systems of more than 256 rows use the wide panel kernel (8 waves, up to 2 rows per
// thread) and the three-GEMM update in order, whatever `update`, `panel` and `la` ask for.  Scratch per chain: Vs N*32 doubles, Ts 32*32 doubles, and Wt holding two blocks of
// wt_stride >= W*32 doubles (batch * wt_stride apart).
// `update`: QR_UPDATE_REG (default) applies the whole block reflector in one launch per panel with
// each 32-column tile held in registers (the first Wt block holds a row-major copy of V);
// QR_UPDATE_GEMM uses three MFMA GEMM launches with tile `tile` for the last one.
// `panel`: QR_PANEL_256 (default) factors each panel with 4 waves per chain, a thread per row;
// QR_PANEL_1024 with 16 waves, four threads per row.
enum QrUpdate : int { QR_UPDATE_GEMM = 0, QR_UPDATE_REG = 2 };
// Look-ahead of the blocked QR (QR_UPDATE_REG) and LU: with s2 set, the trailing update of panel k
// first covers panel k + 1's 32 columns on the main stream; panel k + 1 then runs on the main
// stream beside the rest of the update on s2 (fork and join through the two events, which a graph
// capture of the main stream records as parallel branches).  V2 / T2 / P2 are the parity-1 panel
// buffers (QR: V N*32 and T 32*32 doubles per chain; LU: L, L11^{-1} and the row order N ints).
struct LookAhead {
    hipStream_t s2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    double* V2 = nullptr;
    double* T2 = nullptr;
    int* P2 = nullptr;
    bool fuse = true;   // panel k + 1 applies update k to its own columns (no separate block-0 launch)
};
enum QrPanel : int { QR_PANEL_256 = 0, QR_PANEL_1024 = 1 };
// rhs_identity: columns [N, 2N) are the identity and are not read (QR_UPDATE_REG, N <= 256, W >= 2N).
hipError_t launch_qr_blocked(double* buf, long long stride, int N, int W, int batch, double* Vs, double* Ts,
                             double* Wt, long long wt_stride, int tile, int update, int panel, hipStream_t s,
                             const LookAhead* la = nullptr, bool rhs_identity = false);
// Blocked LU with partial pivoting of the leading N x N block of the same row-major N x W system
// (N <= 256), applied to the W - N trailing columns: on exit the leading block holds U and the
// trailing columns L^{-1} P rhs.  Scratch per chain: Ls N*32 doubles, Li 32*32 doubles, Ps N ints.
// A zero pivot sets HEALTH_ZERO_PIVOT.
hipError_t launch_lu_blocked(double* buf, long long stride, int N, int W, int batch, double* Ls, double* Li, int* Ps,
                             uint32_t* health, hipStream_t s, const LookAhead* la = nullptr);
// Blocked in-place back-substitution of the same system: `block`-row diagonal solves (thread per
// right-hand side; 32, 64 or 128 rows) and MFMA GEMM updates of the rows above (`tile` = GEMM tile).
hipError_t launch_trsm_blocked(double* buf, long long stride, int N, int W, int c0, int P, int batch, int tile,
                               hipStream_t s, int block = 32);
// In-place X = R^{-1} Z, R the upper triangle of the leading N x N block, Z columns
// [c0, c0+P) of the same row-major N x W buffer; one-kernel form (DQMC_TRSM=v1).
hipError_t launch_trsm_upper_v1(double* buf, long long stride, int N, int W, int c0, int P, int batch,
                                hipStream_t s);

// Pre-pivoted UDV input: M column-major (ld N); writes buf[r][j] = M(r, perm[j]) with perm
// sorting the columns of M by descending 2-norm, and buf[r][N+j] = delta_rj (row length W >= 2N).
// `norms` (N per chain), when given, holds the columns' squared norms as the rank kernel would sum
// them, so that M is not read for them.  Without `identity` the delta columns are left unwritten (for
// a QR that takes them as the identity).
hipError_t launch_udv_prepare(const double* M, long long sM, double* buf, long long sbuf, int W, int* perm,
                              int N, int batch, hipStream_t s, const double* norms = nullptr, bool identity = true);
// After QR: D_i = |R_ii|, Tt (column-major N x N) with Tt(perm[j], i) = R_ij / D_i (j >= i),
// U (column-major) = Q read from the trailing block.
hipError_t launch_udv_finish(const double* buf, long long sbuf, int W, const int* perm, double* U, long long sU,
                             double* D, long long sD, double* Tt, long long sTt, int N, uint32_t* health, int batch,
                             hipStream_t s);

// Builds the row-major system for (1 + u diag(d) v^T)^{-1}:
//   buf[i][j] = sl_i u(j,i) + sr_i v(j,i),  buf[i][N+j] = sl_i u(j,i),
// with sl = 1/max(1,d), sr = min(1,d).
hipError_t launch_build_00bb(const double* u, long long su, const double* v, long long sv, const double* d,
                             long long sd, double* buf, long long sbuf, int N, int W, int batch,
                             hipStream_t s);
// buf[r][c0 + j] = alpha * scale(r) * X(j, r) for a column-major X (i.e. the scaled transpose).
hipError_t launch_build_rhs(const double* X, long long sX, DiagArg d, double alpha, double* buf, long long sbuf,
                            int N, int W, int c0, int batch, hipStream_t s);
// Copies block columns [c0, c0+N) of the row-major buffer into a column-major matrix;
// transpose = 0 gives out(j, r) = buf[r][c0+j] (the block is stored transposed), 1 gives
// out(r, j) = buf[r][c0+j].
hipError_t launch_extract(const double* buf, long long sbuf, int W, int c0, double* out, long long so, int N,
                          int transpose, int batch, hipStream_t s);

// Signed-max wrap error (max_ij (Gnew - G)_ij, include/utils/stable_numerics.hpp:42) and max-abs
// difference of a new FP64 G against the working copy; then commits new -> G64 (and G32 if
// non-null).  err_signed / err_abs are per-chain running maxima (updated in place; the caller
// starts them at 0 like src/dqmc_core.cpp:211-212).  Sets HEALTH_NONFINITE_G for a non-finite
// entry on either side and HEALTH_WATCHDOG when *bound > 0 and the max-abs difference exceeds it.
// part: per-chain scratch of CMP_PART_SLOTS x 3 doubles for the two-stage reduction.
#define CMP_PART_SLOTS 16
hipError_t launch_compare_commit(const double* Gnew, double* G64, float* G32, long long sG, int N,
                                 double* err_signed, double* err_abs, const double* bound, uint32_t* health,
                                 double* part, int batch, hipStream_t s);
// The second stage alone: folds np partials per chain (signed max, max-abs, non-finite flag) into the
// running maxima and the health flags, as launch_compare_commit does.
hipError_t launch_compare_fold(const double* part, int np, double* err_signed, double* err_abs, const double* bound,
                               uint32_t* health, int batch, hipStream_t s);
// Same but the working matrix is FP64 only (dynamic G(t,0), G(0,t)); no commit.
hipError_t launch_compare_only(const double* Gnew, const double* Gcur, long long sG, int N, double* err_signed,
                               uint32_t* health, int batch, hipStream_t s);
// HEALTH_NONFINITE_G for chains whose N x N matrix holds a non-finite entry.
hipError_t launch_finite_check(const double* A, long long sA, int N, uint32_t* health, int batch, hipStream_t s);

// Rank-1 Sherman-Morrison update of one chain's G (N x N column-major) at site i for the field
// flip with delta = e^{-2 alpha s} - 1 (src/hubbard.cpp:132-139).
hipError_t launch_site_update(double* G, int N, int i, double delta, hipStream_t s);

// Elementwise helpers (n elements per chain).
hipError_t launch_set_identity(double* A, long long s, int N, double diagval, int batch, hipStream_t st);
hipError_t launch_copy_d(const double* src, long long ss, double* dst, long long sd, long long n, int batch,
                         hipStream_t st);
hipError_t launch_copy_f2d(const float* src, long long ss, double* dst, long long sd, long long n, int batch,
                           hipStream_t st);
hipError_t launch_copy_d2f(const double* src, long long ss, float* dst, long long sd, long long n, int batch,
                           hipStream_t st);
// dst = a*src + b*I (N x N)
hipError_t launch_axpi(const double* src, long long ss, double* dst, long long sd, double a, double b, int N,
                       int batch, hipStream_t st);
hipError_t launch_fill_d(double* dst, long long n, double v, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* dst, long long n, uint32_t v, hipStream_t st);

// ---- observables on the device (dqmc_measure.hip) ------------------------------------------
// Equal-time estimators of all L stored slices folded into cache[b][0..2] (filling, double
// occupancy, s-wave pairing); part is [batch][L][3] scratch.
hipError_t launch_measure_equaltime(const double* vec, long long sv, int L, int N, double* part, double* cache,
                                    int cstride, int batch, hipStream_t s);
// Dynamic estimators (flags: 1 G(k,t), 2 LDOS, 4 chi_perp) of all L slices added to
// cache[b][3 ..]; costab [nk][N] holds cos(k . r_d).
hipError_t launch_measure_dynamic(const double* gt0, const double* g0t, long long sv, int L, int nl,
                                  const double* costab, int nk, double* cache, int cstride, int flags, int batch,
                                  hipStream_t s);
// data[b][bin] = cache[b] / count (eq_count for the first neq entries, dyn_count after), cache = 0.
hipError_t launch_measure_push(double* cache, int cstride, double* data, long long dstride, int bin, int neq,
                               long long eq_count, long long dyn_count, int batch, hipStream_t s);

// ---- parameter adjoint at fixed fields (dqmc_adjoint.hip) ----------------------------------
struct AdjointArgs {
    const double* gf0t;         // [batch][L] N x N: G(tau, tau+t') of the rolled pass
    const double* gft0;         // [batch][L] N x N: G(tau+t', tau)
    const double* gftt;         // [batch][L] N x N: slice 0 holds G(tau)
    long long sv;               // chain stride of the three (L*N*N)
    const double* gbar_diag;    // [batch][L][N] cotangent of G(t)_ii (original slice order)
    const double* gbar_logdet;  // [batch] cotangent of log|det(1 + B_L..B_1)|, may be null
    const int8_t* fields;       // rolled fields [batch][L][N]
    const int* nb;              // [N][4] neighbours (x+1, y+1, x-1, y-1)
    double* partial;            // [batch][L+1][3] scratch
    double* grad;               // [batch][3] accumulated d/d(nnt, mu, u)
    int N, L, tau;
    double alpha, dt, dalpha;   // dalpha = d alpha / d u
};
// Contracts one rolled pass into grad (two launches: per-slice partials, fixed-order sum).
hipError_t launch_adjoint_contract(const AdjointArgs& a, int batch, hipStream_t s);
// dst[b][k][i] = src[b][(k + shift) % L][i]
hipError_t launch_roll_fields(const int8_t* src, int8_t* dst, int L, int N, int shift, int batch, hipStream_t s);
// log|det(1 + u diag(d) v^T)| per chain from the factored launch_build_00bb system.
hipError_t launch_logdet_00bb(const double* buf, long long sbuf, int W, const double* d, long long sd, int N,
                              double* out, int batch, hipStream_t s);
// out[b][t][i] = vec_b[t](i, i)
hipError_t launch_gf_diag(const double* vec, long long sv, int L, int N, double* out, int batch, hipStream_t s);

}  // namespace dqmc

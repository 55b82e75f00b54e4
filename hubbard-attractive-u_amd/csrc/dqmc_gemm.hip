// dqmc_gemm.hip -- batched GEMM on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32) with fused
// diagonal scalings (slice wraps, B-chain, UDV and recompose products), and the Kronecker-factored
// hopping exponential on MFMA.
#include "dqmc_device.h"

#include <type_traits>

namespace dqmc {

// ---------------------------------------------------------------------------------------------
//  GEMM.  256 threads = 4 waves in a 2 x 2 grid; each wave owns a (BM/2) x (BN/2) block of
//  16 x 16 MFMA tiles.  K advances in steps of BK through LDS (k-major tiles so that the
//  fragment reads of consecutive lanes hit consecutive LDS words).
// ---------------------------------------------------------------------------------------------
template <typename T, int BM, int BN, int BK>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs<T> g)
{
    constexpr int WM = BM / 2, WN = BN / 2, TM = WM / 16, TN = WN / 16;
    constexpr int PAD = (sizeof(T) == 8) ? 1 : 2;
    __shared__ T As[BK][BM + PAD];
    __shared__ T Bs[BK][BN + PAD];
    typedef typename Mfma<T>::acc_t acc_t;

    const int gm = (g.M + BM - 1) / BM, gn = (g.Nc + BN - 1) / BN;
    int b, t;
    chain_tile(gm * gn, g.batch, b, t);
    const int m0 = (t % gm) * BM, n0 = (t / gm) * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const T* __restrict__ A = g.A + (long long)b * g.sA;
    const T* __restrict__ B = g.B + (long long)b * g.sB;
    T* __restrict__ C = g.C + (long long)b * g.sC;

    acc_t acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // The next K tile is fetched into registers while the current one is multiplied from LDS.
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;
    T ra[NA], rb[NB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int e = u * 256 + tid;
            int i, kk;
            if (!g.tA) { i = e % BM; kk = e / BM; } else { kk = e % BK; i = e / BK; }
            const int gi = m0 + i, gk = k0 + kk;
            T v = T(0);
            if (gi < g.M && gk < g.K) {
                v = g.tA ? A[gk + (long long)gi * g.lda] : A[gi + (long long)gk * g.lda];
                if (g.innd.mode != DIAG_NONE) v = T((double)v * diag_value(g.innd, b, gk));
            }
            ra[u] = v;
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int e = u * 256 + tid;
            int j, kk;
            if (!g.tB) { kk = e % BK; j = e / BK; } else { j = e % BN; kk = e / BN; }
            const int gj = n0 + j, gk = k0 + kk;
            T v = T(0);
            if (gj < g.Nc && gk < g.K) v = g.tB ? B[gj + (long long)gk * g.ldb] : B[gk + (long long)gj * g.ldb];
            rb[u] = v;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < g.K; k0 += BK) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int e = u * 256 + tid;
            int i, kk;
            if (!g.tA) { i = e % BM; kk = e / BM; } else { kk = e % BK; i = e / BK; }
            As[kk][i] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int e = u * 256 + tid;
            int j, kk;
            if (!g.tB) { kk = e % BK; j = e / BK; } else { j = e % BN; kk = e / BN; }
            Bs[kk][j] = rb[u];
        }
        __syncthreads();
        if (k0 + BK < g.K) fetch(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T af[TM], bf[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) af[i] = As[kk + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < TN; ++j) bf[j] = Bs[kk + (lane >> 4)][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = Mfma<T>::mma(af[i], bf[j], acc[i][j]);
        }
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm * WM + i * 16 + Mfma<T>::row(lane, r);
                const int gj = n0 + wn * WN + j * 16 + (lane & 15);
                if (gi < g.M && gj < g.Nc) {
                    double v = (double)acc[i][j][r] * g.alpha;
                    if (g.rowd.mode != DIAG_NONE) v *= diag_value(g.rowd, b, gi);
                    if (g.cold.mode != DIAG_NONE) v *= diag_value(g.cold, b, gj);
                    const long long off = g.tC ? (gj + (long long)gi * g.ldc) : (gi + (long long)gj * g.ldc);
                    if (g.beta != 0.0) v += g.beta * (double)C[off];
                    C[off] = T(v);
                }
            }
}

// 64 x 64 output tile with compile-time operand layouts, LDS tiles double-buffered (one barrier
// per K step) and the global loads two K steps ahead in two register stages, so a load has the
// MFMA work of two steps to arrive.  The loop runs two steps per iteration so that every register
// stage is addressed statically.
template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm64_kernel(GemmArgs<T> g)
{
    constexpr int BM = 64, BN = 64, BK = 16, WM = 32, WN = 32, NA = BM * BK / 256, NB = BN * BK / 256;
    constexpr int PAD = (sizeof(T) == 8) ? 1 : 2;
    __shared__ T As[2][BK][BM + PAD];
    __shared__ T Bs[2][BK][BN + PAD];
    __shared__ double sk[256], srow[BM], scol[BN];   // diagonal scalings, evaluated once per tile
    typedef typename Mfma<T>::acc_t acc_t;

    const int gm = (g.M + BM - 1) / BM, gn = (g.Nc + BN - 1) / BN;
    int b, t;
    chain_tile(gm * gn, g.batch, b, t);
    const int m0 = (t % gm) * BM, n0 = (t / gm) * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const T* __restrict__ A = g.A + (long long)b * g.sA;
    const T* __restrict__ B = g.B + (long long)b * g.sB;
    T* __restrict__ C = g.C + (long long)b * g.sC;
    const bool innd = g.innd.mode != DIAG_NONE;
    if (innd)
        for (int kk = tid; kk < g.K && kk < 256; kk += 256) sk[kk] = diag_value(g.innd, b, kk);
    if (tid < BM) srow[tid] = (g.rowd.mode != DIAG_NONE && m0 + tid < g.M) ? diag_value(g.rowd, b, m0 + tid) : 1.0;
    else if (tid < BM + BN)
        scol[tid - BM] = (g.cold.mode != DIAG_NONE && n0 + tid - BM < g.Nc) ? diag_value(g.cold, b, n0 + tid - BM) : 1.0;
    __syncthreads();

    acc_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    T ra0[NA], rb0[NB], ra1[NA], rb1[NB];
    auto fetch = [&](int k0, T (&ra)[NA], T (&rb)[NB]) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int e = u * 256 + tid;
            const int i = TA ? e / BK : e % BM, kk = TA ? e % BK : e / BM;
            const int gi = m0 + i, gk = k0 + kk;
            T v = T(0);
            if (gi < g.M && gk < g.K) v = TA ? A[gk + (long long)gi * g.lda] : A[gi + (long long)gk * g.lda];
            ra[u] = v;   // the inner scaling is applied at the LDS store, so the load is not waited on here
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int e = u * 256 + tid;
            const int j = TB ? e % BN : e / BK, kk = TB ? e / BN : e % BK;
            const int gj = n0 + j, gk = k0 + kk;
            rb[u] = (gj < g.Nc && gk < g.K) ? (TB ? B[gj + (long long)gk * g.ldb] : B[gk + (long long)gj * g.ldb]) : T(0);
        }
    };
    auto stash = [&](int buf, int k0, const T (&ra)[NA], const T (&rb)[NB]) {
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int e = u * 256 + tid;
            const int kk = TA ? e % BK : e / BM;
            T v = ra[u];
            if (innd && k0 + kk < g.K) v = T((double)v * sk[k0 + kk]);
            As[buf][kk][TA ? e / BK : e % BM] = v;
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int e = u * 256 + tid;
            Bs[buf][TB ? e / BN : e % BK][TB ? e % BN : e / BK] = rb[u];
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T af[2], bf[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) af[i] = As[buf][kk + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 2; ++j) bf[j] = Bs[buf][kk + (lane >> 4)][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<T>::mma(af[i], bf[j], acc[i][j]);
        }
    };

    const int nk = (g.K + BK - 1) / BK;
    fetch(0, ra0, rb0);
    if (nk > 1) fetch(BK, ra1, rb1);
    stash(0, 0, ra0, rb0);
    __syncthreads();
    // invariant at the top of step ks: LDS buffer ks&1 holds step ks, register stage (ks+1)&1 holds
    // step ks+1 (if any), register stage ks&1 is free
    for (int ks = 0; ks < nk; ks += 2) {
        if (ks + 2 < nk) fetch((ks + 2) * BK, ra0, rb0);
        compute(0);
        if (ks + 1 < nk) stash(1, (ks + 1) * BK, ra1, rb1);
        __syncthreads();
        if (ks + 1 >= nk) break;
        if (ks + 3 < nk) fetch((ks + 3) * BK, ra1, rb1);
        compute(1);
        if (ks + 2 < nk) stash(0, (ks + 2) * BK, ra0, rb0);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm * WM + i * 16 + Mfma<T>::row(lane, r);
                const int gj = n0 + wn * WN + j * 16 + (lane & 15);
                if (gi < g.M && gj < g.Nc) {
                    double v = (double)acc[i][j][r] * g.alpha;
                    if (g.rowd.mode != DIAG_NONE) v *= srow[gi - m0];
                    if (g.cold.mode != DIAG_NONE) v *= scol[gj - n0];
                    const long long off = g.tC ? (gj + (long long)gi * g.ldc) : (gi + (long long)gj * g.ldc);
                    if (g.beta != 0.0) v += g.beta * (double)C[off];
                    C[off] = T(v);
                }
            }
}

// FP64 128 x 128 output tile: 4 waves in a 2 x 2 grid, each a 64 x 64 block of 4 x 4 MFMA tiles,
// so every LDS fragment feeds four MFMAs and every global byte eight times the flops of the 64 x 64
// tile.  K advances by 8 through double-buffered LDS, with the global loads of the next step in one
// register stage, so that two workgroups fit a CU.
// With PAIR the tile is the sum of two products that share the shape and transposes,
//   C = diag(r1) op(A) op(B) diag(c1) + diag(r2) op(A2) op(B2) diag(c2)
// (g.rowd, g.cold and the pair's rowd2, cold2), as one K loop over both: the row scalings are applied to
// the A operands and the column scalings to the B operands as they are stored in LDS.  When p.R is
// set, the tiles of column block 0 also store their scaled first A operand, diag(r1) op(A), row-major
// at p.R (row i at p.R + i * p.ldr).
// With COMMIT the tile is stored over the previous contents of C after comparing with them: its signed
// maximum and maximum absolute difference (new - old) and a non-finite flag go to
// part[(chain * tiles + tile) * 3 + {0, 1, 2}], to be folded per chain by launch_compare_fold.
enum { G128_PLAIN = 0, G128_PAIR = 1, G128_COMMIT = 2 };
template <bool TA, bool TB, int MODE>
__device__ __forceinline__ void gemm128_body(const GemmArgs<double>& g, const GemmPair& p, double* __restrict__ part)
{
    constexpr int BM = 128, BN = 128, BK = 8, WM = 64, WN = 64, NA = BM * BK / 256, NB = BN * BK / 256;
    constexpr bool PAIR = MODE == G128_PAIR, COMMIT = MODE == G128_COMMIT;
    constexpr int NP = PAIR ? 2 : 1;
    __shared__ double As[2][BK][BM + 1];
    __shared__ double Bs[2][BK][BN + 1];
    // diagonal scalings, evaluated once per tile (PAIR: per product, applied to the operands)
    __shared__ double sk[256], srow[NP][BM], scol[NP][BN];
    typedef Mfma<double>::acc_t acc_t;

    const int gm = (g.M + BM - 1) / BM, gn = (g.Nc + BN - 1) / BN;
    int b, t;
    chain_tile(gm * gn, g.batch, b, t);
    const int m0 = (t % gm) * BM, n0 = (t / gm) * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const double* __restrict__ A = g.A + (long long)b * g.sA;
    const double* __restrict__ B = g.B + (long long)b * g.sB;
    const double* __restrict__ A2 = PAIR ? p.A2 + (long long)b * p.sA2 : nullptr;
    const double* __restrict__ B2 = PAIR ? p.B2 + (long long)b * p.sB2 : nullptr;
    double* __restrict__ R = (PAIR && p.R && n0 == 0) ? p.R + (long long)b * p.sR : nullptr;
    double* __restrict__ C = g.C + (long long)b * g.sC;
    const bool innd = !PAIR && g.innd.mode != DIAG_NONE;
    if (innd)
        for (int kk = tid; kk < g.K && kk < 256; kk += 256) sk[kk] = diag_value(g.innd, b, kk);
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const DiagArg& rd = q ? p.rowd2 : g.rowd;
        const DiagArg& cd = q ? p.cold2 : g.cold;
        if (tid < BM) srow[q][tid] = (rd.mode != DIAG_NONE && m0 + tid < g.M) ? diag_value(rd, b, m0 + tid) : 1.0;
        else scol[q][tid - BM] = (cd.mode != DIAG_NONE && n0 + tid - BM < g.Nc) ? diag_value(cd, b, n0 + tid - BM) : 1.0;
    }
    __syncthreads();

    acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    // step ks covers k0 = (ks mod nk1) * BK of product ks / nk1
    const int nk1 = (g.K + BK - 1) / BK;
    auto fetch = [&](int ks, double (&ra)[NA], double (&rb)[NB]) {
        const bool second = PAIR && ks >= nk1;
        const double* __restrict__ Aq = second ? A2 : A;
        const double* __restrict__ Bq = second ? B2 : B;
        const int lda = second ? p.lda2 : g.lda, ldb = second ? p.ldb2 : g.ldb;
        const int k0 = (second ? ks - nk1 : ks) * BK;
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int e = u * 256 + tid;
            const int i = TA ? e / BK : e % BM, kk = TA ? e % BK : e / BM;
            const int gi = m0 + i, gk = k0 + kk;
            ra[u] = (gi < g.M && gk < g.K) ? (TA ? Aq[gk + (long long)gi * lda] : Aq[gi + (long long)gk * lda]) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int e = u * 256 + tid;
            const int j = TB ? e % BN : e / BK, kk = TB ? e / BN : e % BK;
            const int gj = n0 + j, gk = k0 + kk;
            rb[u] = (gj < g.Nc && gk < g.K) ? (TB ? Bq[gj + (long long)gk * ldb] : Bq[gk + (long long)gj * ldb]) : 0.0;
        }
    };
    auto stash = [&](int buf, int ks, const double (&ra)[NA], const double (&rb)[NB]) {
        const int q = (PAIR && ks >= nk1) ? 1 : 0;
        const int k0 = (q ? ks - nk1 : ks) * BK;
#pragma unroll
        for (int u = 0; u < NA; ++u) {
            const int e = u * 256 + tid;
            const int kk = TA ? e % BK : e / BM, i = TA ? e / BK : e % BM;
            double v = ra[u];
            if (innd && k0 + kk < g.K) v *= sk[k0 + kk];
            if (PAIR) {
                v *= srow[q][i];
                if (R && q == 0 && m0 + i < g.M && k0 + kk < g.K) R[(long long)(m0 + i) * p.ldr + k0 + kk] = v;
            }
            As[buf][kk][i] = v;
        }
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            const int e = u * 256 + tid;
            const int j = TB ? e % BN : e / BK;
            double v = rb[u];
            if (PAIR) v *= scol[q][j];
            Bs[buf][TB ? e / BN : e % BK][j] = v;
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = As[buf][kk + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[j] = Bs[buf][kk + (lane >> 4)][wn * WN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<double>::mma(af[i], bf[j], acc[i][j]);
        }
    };

    const int nk = NP * nk1;
    double ra[NA], rb[NB];
    fetch(0, ra, rb);
    stash(0, 0, ra, rb);
    __syncthreads();
    // one register stage: step ks + 1 is loaded while step ks is multiplied from LDS buffer ks & 1
    for (int ks = 0; ks < nk; ++ks) {
        if (ks + 1 < nk) fetch(ks + 1, ra, rb);
        compute(ks & 1);
        if (ks + 1 < nk) stash((ks + 1) & 1, ks + 1, ra, rb);
        __syncthreads();
    }

    double es = -DBL_MAX, ea = 0.0;
    bool bad = false;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        // the previous values of this row of MFMA tiles (COMMIT, or beta != 0), loaded together before any
        // of its stores: a load after a store to the same array would otherwise wait for it
        double prev[4][4];
        if (COMMIT || g.beta != 0.0) {
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int gi = m0 + wm * WM + i * 16 + Mfma<double>::row(lane, r);
                    const int gj = n0 + wn * WN + j * 16 + (lane & 15);
                    const long long off = g.tC ? (gj + (long long)gi * g.ldc) : (gi + (long long)gj * g.ldc);
                    prev[j][r] = (gi < g.M && gj < g.Nc) ? C[off] : 0.0;
                }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = m0 + wm * WM + i * 16 + Mfma<double>::row(lane, r);
                const int gj = n0 + wn * WN + j * 16 + (lane & 15);
                if (gi < g.M && gj < g.Nc) {
                    double v = acc[i][j][r] * g.alpha;
                    if (!PAIR && g.rowd.mode != DIAG_NONE) v *= srow[0][gi - m0];
                    if (!PAIR && g.cold.mode != DIAG_NONE) v *= scol[0][gj - n0];
                    const long long off = g.tC ? (gj + (long long)gi * g.ldc) : (gi + (long long)gj * g.ldc);
                    if (COMMIT) {
                        const double old = prev[j][r], d = v - old;
                        es = fmax(es, d);
                        ea = fmax(ea, fabs(d));
                        bad |= !isfinite(v) || !isfinite(old);
                    } else if (g.beta != 0.0) {
                        v += g.beta * prev[j][r];
                    }
                    C[off] = v;
                }
            }
    }
    if constexpr (COMMIT) {
        // the tile's maxima over its four waves (max is exact: the order of the pairings is free)
        __shared__ double red[3][4];
        double fl = bad ? 1.0 : 0.0;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            es = fmax(es, __shfl_xor(es, o, 64));
            ea = fmax(ea, __shfl_xor(ea, o, 64));
            fl = fmax(fl, __shfl_xor(fl, o, 64));
        }
        if (lane == 0) { red[0][wave] = es; red[1][wave] = ea; red[2][wave] = fl; }
        __syncthreads();
        if (tid < 3) {
            const double m = fmax(fmax(red[tid][0], red[tid][1]), fmax(red[tid][2], red[tid][3]));
            part[((long long)b * gm * gn + t) * 3 + tid] = m;
        }
    }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm128_kernel(GemmArgs<double> g)
{
    gemm128_body<TA, TB, G128_PLAIN>(g, GemmPair(), nullptr);
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm128_pair_kernel(GemmArgs<double> g,
                                                                                                  GemmPair p)
{
    gemm128_body<TA, TB, G128_PAIR>(g, p, nullptr);
}

template <bool TA, bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm128_commit_kernel(GemmArgs<double> g,
                                                                                                    double* part)
{
    gemm128_body<TA, TB, G128_COMMIT>(g, GemmPair(), part);
}

template <bool TA, bool TB>
static hipError_t launch_gemm128(const GemmArgs<double>& a, hipStream_t s)
{
    const long long blocks = (long long)((a.M + 127) / 128) * ((a.Nc + 127) / 128) * a.batch;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm128_kernel<TA, TB>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_gemm_pair(const GemmArgs<double>& a, const GemmPair& p, hipStream_t s)
{
    if (a.batch <= 0 || a.M <= 0 || a.Nc <= 0) return hipSuccess;
    if (a.M < 128 || a.Nc < 128 || a.K < 64 || a.innd.mode != DIAG_NONE || a.beta != 0.0 || a.tA != 1 || a.tB != 0)
        return hipErrorInvalidValue;
    const long long blocks = (long long)((a.M + 127) / 128) * ((a.Nc + 127) / 128) * a.batch;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm128_pair_kernel<true, false>), dim3((unsigned)blocks), dim3(256), 0, s, a, p);
    return hipGetLastError();
}

hipError_t launch_gemm_commit(const GemmArgs<double>& a, double* part, int max_tiles, hipStream_t s)
{
    if (a.batch <= 0 || a.M <= 0 || a.Nc <= 0) return hipSuccess;
    if (a.M < 128 || a.Nc < 128 || a.K < 64 || (a.innd.mode != DIAG_NONE && a.K > 256) || a.beta != 0.0 ||
        a.tA != 0 || a.tB != 1)
        return hipErrorInvalidValue;
    const long long tiles = (long long)((a.M + 127) / 128) * ((a.Nc + 127) / 128);
    if (tiles > max_tiles || tiles * a.batch > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm128_commit_kernel<false, true>), dim3((unsigned)(tiles * a.batch)), dim3(256), 0, s, a, part);
    return hipGetLastError();
}

template <typename T, bool TA, bool TB>
static hipError_t launch_gemm64(const GemmArgs<T>& a, hipStream_t s)
{
    const long long blocks = (long long)((a.M + 63) / 64) * ((a.Nc + 63) / 64) * a.batch;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm64_kernel<T, TA, TB>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T, int BM, int BN>
static hipError_t launch_gemm_tile(const GemmArgs<T>& a, hipStream_t s)
{
    const long long blocks = (long long)((a.M + BM - 1) / BM) * ((a.Nc + BN - 1) / BN) * a.batch;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, 16>), dim3((unsigned)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_gemm(const GemmArgs<T>& a, int tile, hipStream_t s)
{
    if (a.batch <= 0 || a.M <= 0 || a.Nc <= 0) return hipSuccess;
    if (tile == TILE_128x128) {
        // FP64 products with both output sides >= 128, K >= 64 and an inner scaling table that fits;
        // other shapes (and FP32) take the 64 x 64 prefetching tile
        if constexpr (std::is_same<T, double>::value) {
            if (a.M >= 128 && a.Nc >= 128 && a.K >= 64 && (a.innd.mode == DIAG_NONE || a.K <= 256)) {
                if (a.tA) return a.tB ? launch_gemm128<true, true>(a, s) : launch_gemm128<true, false>(a, s);
                return a.tB ? launch_gemm128<false, true>(a, s) : launch_gemm128<false, false>(a, s);
            }
        }
        tile = TILE_64x64_PF;
    }
    switch (tile) {
        case TILE_32x32: return launch_gemm_tile<T, 32, 32>(a, s);
        case TILE_64x32: return launch_gemm_tile<T, 64, 32>(a, s);
        case TILE_64x64_PF:
            if (a.innd.mode != DIAG_NONE && a.K > 256) return launch_gemm_tile<T, 64, 64>(a, s);
            if (a.tA) return a.tB ? launch_gemm64<T, true, true>(a, s) : launch_gemm64<T, true, false>(a, s);
            return a.tB ? launch_gemm64<T, false, true>(a, s) : launch_gemm64<T, false, false>(a, s);
        default: return launch_gemm_tile<T, 64, 64>(a, s);
    }
}
template hipError_t launch_gemm<double>(const GemmArgs<double>&, int, hipStream_t);
template hipError_t launch_gemm<float>(const GemmArgs<float>&, int, hipStream_t);

// ---------------------------------------------------------------------------------------------
//  Kronecker-factored hopping exponential on MFMA.
//  A vector v of the lattice (site i = x + nl*y) is an nl x nl tile V(x, y).  (E(x)E) v is the
//  tile E V E^T = E V E (E symmetric).  With 16x16x4 MFMA fragments:
//    step 1  Y^T = V^T E^T : A-fragment = V^T read from memory, B-fragment = E (constant);
//    step 2  Z   = Y E     : the A-fragment of step k is exactly item k of step 1's accumulator
//  (the f64 accumulator layout puts Y^T(4k + l/16, l%16) in lane l, item k), so the tile never
//  leaves registers.  8 MFMAs per lattice vector.
// ---------------------------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ int kron_row(int lane, int r) { return Mfma<T>::row(lane, r); }

// Transform one tile: vin(x', y') supplied by `ld(x', y')`, result passed to `st(x, y, value)`.
template <typename T, typename LD, typename ST>
__device__ __forceinline__ void kron_tile(const T (&ef)[4], const T (&ef2)[4], int lane, int nl, LD ld, ST st)
{
    typedef typename Mfma<T>::acc_t acc_t;
    acc_t y = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int xp = (lane >> 4) + 4 * s, yp = lane & 15;   // A(row = y', k = x') = V(x', y')
        const T v = (xp < nl && yp < nl) ? ld(xp, yp) : T(0);
        y = Mfma<T>::mma(v, ef[s], y);                          // B(k = x', col = x) = E[x][x']
    }
    acc_t z = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; ++s) z = Mfma<T>::mma(y[s], ef2[s], z);   // A(row = x, k = y') = Y(x, y')
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int x = kron_row<T>(lane, r), yy = lane & 15;
        if (x < nl && yy < nl) st(x, yy, z[r]);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void kron_left_kernel(KronArgs<T> a)
{
    const int b = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nl = a.nl, N = a.N;
    T ef[4], ef2[4];
    load_efrag<T>(a.E, lane, ef, ef2);
    const T* X = a.X + (long long)b * a.sX;
    T* C = a.C + (long long)b * a.sC;
    const int cbase = blockIdx.x * 16 + wave * 4;
    // The diagonal scalings depend on the lattice site only, so each lane evaluates the ones of
    // its four input sites and four output sites once for all columns.
    double din[4], dout[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int xp = (lane >> 4) + 4 * s, yp = lane & 15;
        din[s] = (a.innd.mode != DIAG_NONE && xp < nl && yp < nl) ? diag_value(a.innd, b, xp + nl * yp) : 1.0;
        const int x = kron_row<T>(lane, s), y = lane & 15;
        dout[s] = (a.rowd.mode != DIAG_NONE && x < nl && y < nl) ? diag_value(a.rowd, b, x + nl * y) : 1.0;
    }
    // all four lattice vectors of the wave are loaded before the first MFMA
    T vin[4][4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        const int c = cbase + cc;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int xp = (lane >> 4) + 4 * s, yp = lane & 15;
            T v = T(0);
            if (c < N && xp < nl && yp < nl) {
                v = X[(long long)c * N + xp + nl * yp];
                if (a.innd.mode != DIAG_NONE) v = T((double)v * din[s]);
            }
            vin[cc][s] = v;
        }
    }
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        const int c = cbase + cc;
        if (c >= N) break;
        const double cs = (a.cold.mode != DIAG_NONE) ? diag_value(a.cold, b, c) : 1.0;
        T* cc_out = C + (long long)c * N;
        typedef typename Mfma<T>::acc_t acc_t;
        acc_t y = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) y = Mfma<T>::mma(vin[cc][s], ef[s], y);
        acc_t z = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) z = Mfma<T>::mma(y[s], ef2[s], z);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int x = kron_row<T>(lane, r), yy = lane & 15;
            if (x < nl && yy < nl) cc_out[x + nl * yy] = T((double)z[r] * cs * dout[r]);
        }
    }
}

// Right multiplication: 16 consecutive rows per workgroup are staged through LDS (coalesced
// column segments), transformed per row, and written back.
template <typename T>
__global__ __launch_bounds__(256) void kron_right_kernel(KronArgs<T> a)
{
    __shared__ T tile[256][17];
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nl = a.nl, N = a.N;
    const int r0 = blockIdx.x * 16;
    const T* X = a.X + (long long)b * a.sX;
    T* C = a.C + (long long)b * a.sC;
    for (int e = tid; e < N * 16; e += 256) {
        const int rr = e & 15, c = e >> 4;
        const int row = r0 + rr;
        T v = T(0);
        if (row < N) {
            v = X[row + (long long)c * N];
            if (a.innd.mode != DIAG_NONE) v = T((double)v * diag_value(a.innd, b, c));
        }
        tile[c][rr] = v;
    }
    __syncthreads();
    T ef[4], ef2[4];
    load_efrag<T>(a.E, lane, ef, ef2);
    for (int rr = wave; rr < 16; rr += 4) {
        T vin[4];
        // read the whole row first: the tile is transformed in place
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int xp = (lane >> 4) + 4 * s, yp = lane & 15;
            vin[s] = (xp < nl && yp < nl) ? tile[xp + nl * yp][rr] : T(0);
        }
        kron_tile<T>(ef, ef2, lane, nl,
            [&](int xp, int yp) {
                return vin[(xp - (lane >> 4)) >> 2];
            },
            [&](int x, int y, T z) { tile[x + nl * y][rr] = z; });
    }
    __syncthreads();
    for (int e = tid; e < N * 16; e += 256) {
        const int rr = e & 15, c = e >> 4;
        const int row = r0 + rr;
        if (row < N) {
            double v = (double)tile[c][rr];
            if (a.rowd.mode != DIAG_NONE) v *= diag_value(a.rowd, b, row);
            if (a.cold.mode != DIAG_NONE) v *= diag_value(a.cold, b, c);
            C[row + (long long)c * N] = T(v);
        }
    }
}

// Left product on the 16 x 16 lattice: each wave transforms 4 columns of X, one after the other
// through its own transpose buffer.
template <typename T>
__global__ __launch_bounds__(256) void kron_left16_kernel(KronArgs<T> a)
{
    __shared__ T zt[4][16][17];   // per wave: Z(i, j) at [j][i]
    const int b = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T ef[4], ef2[4];
    load_efrag<T>(a.E, lane, ef, ef2);
    const T* X = a.X + (long long)b * a.sX;
    T* C = a.C + (long long)b * a.sC;
    const int cbase = blockIdx.x * 16 + wave * 4;
    double din[4], dout[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int site = lane + 64 * s;
        din[s] = (a.innd.mode != DIAG_NONE) ? diag_value(a.innd, b, site) : 1.0;
        dout[s] = (a.rowd.mode != DIAG_NONE) ? diag_value(a.rowd, b, site) : 1.0;
    }
    T vin[4][4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const T v = X[(long long)(cbase + cc) * 256 + lane + 64 * s];
            vin[cc][s] = (a.innd.mode != DIAG_NONE) ? T((double)v * din[s]) : v;
        }
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
        typename Mfma<T>::acc_t z;
        kron16_vec<T>(ef, ef2, vin[cc], z);
#pragma unroll
        for (int r = 0; r < 4; ++r) zt[wave][lane & 15][Mfma<T>::row(lane, r)] = z[r];
        wave_lds_sync();
        const int c = cbase + cc;
        const double cs = (a.cold.mode != DIAG_NONE) ? diag_value(a.cold, b, c) : 1.0;
        T zv[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) zv[s] = zt[wave][(lane >> 4) + 4 * s][lane & 15];
        wave_lds_sync();   // the buffer is read before the next column overwrites it
#pragma unroll
        for (int s = 0; s < 4; ++s)   // site l + 64 s = i + 16 j with i = l%16, j = l/16 + 4 s
            C[(long long)c * 256 + lane + 64 * s] = T((double)zv[s] * cs * dout[s]);
    }
}

// Right product on the 16 x 16 lattice: 32 rows of X per workgroup (every column read and written
// as one 256-byte segment, two rows per 16-byte access), staged row-major in LDS with row stride
// 273 and one pad word per 16 sites so that both the lattice-vector reads and the transposed
// accumulator writes spread over the banks; wave w of 8 transforms rows w, w+8, w+16, w+24.
#define KR16_LD 273
#define KR16_ROWS 32
__device__ __forceinline__ int kr16_pad(int site) { return site + (site >> 4); }
template <typename T>
__global__ __launch_bounds__(512) void kron_right16_kernel(KronArgs<T> a)
{
    __shared__ T tile[KR16_ROWS * KR16_LD];
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blockIdx.x * KR16_ROWS;
    const T* X = a.X + (long long)b * a.sX;
    T* C = a.C + (long long)b * a.sC;
#pragma unroll 4
    for (int e = tid; e < KR16_ROWS / 2 * 256; e += 512) {
        const int rr = 2 * (e & 15), c = e >> 4;
        const T* src = X + r0 + rr + (long long)c * 256;
        T v0, v1;
        if (sizeof(T) == 8) {
            const double2 v = *reinterpret_cast<const double2*>(src);
            v0 = T(v.x); v1 = T(v.y);
        } else {
            const float2 v = *reinterpret_cast<const float2*>(src);
            v0 = T(v.x); v1 = T(v.y);
        }
        if (a.innd.mode != DIAG_NONE) {
            const double d = diag_value(a.innd, b, c);
            v0 = T((double)v0 * d); v1 = T((double)v1 * d);
        }
        tile[rr * KR16_LD + kr16_pad(c)] = v0;
        tile[(rr + 1) * KR16_LD + kr16_pad(c)] = v1;
    }
    __syncthreads();
    T ef[4], ef2[4];
    load_efrag<T>(a.E, lane, ef, ef2);
#pragma unroll
    for (int q = 0; q < KR16_ROWS / 8; ++q) {
        T* row = tile + (wave + 8 * q) * KR16_LD;
        T vin[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) vin[s] = row[kr16_pad(lane + 64 * s)];
        typename Mfma<T>::acc_t z;
        kron16_vec<T>(ef, ef2, vin, z);
#pragma unroll
        for (int r = 0; r < 4; ++r) row[kr16_pad(Mfma<T>::row(lane, r) + 16 * (lane & 15))] = z[r];
    }
    __syncthreads();
#pragma unroll 4
    for (int e = tid; e < KR16_ROWS / 2 * 256; e += 512) {
        const int rr = 2 * (e & 15), c = e >> 4;
        double v0 = (double)tile[rr * KR16_LD + kr16_pad(c)], v1 = (double)tile[(rr + 1) * KR16_LD + kr16_pad(c)];
        if (a.rowd.mode != DIAG_NONE) {
            v0 *= diag_value(a.rowd, b, r0 + rr);
            v1 *= diag_value(a.rowd, b, r0 + rr + 1);
        }
        if (a.cold.mode != DIAG_NONE) {
            const double d = diag_value(a.cold, b, c);
            v0 *= d; v1 *= d;
        }
        T* dst = C + r0 + rr + (long long)c * 256;
        if (sizeof(T) == 8) {
            double2 v; v.x = v0; v.y = v1;
            *reinterpret_cast<double2*>(dst) = v;
        } else {
            float2 v; v.x = (float)v0; v.y = (float)v1;
            *reinterpret_cast<float2*>(dst) = v;
        }
    }
}

// Several B-chain factors on the 16 x 16 lattice in one pass (FP64).  The f64 accumulator item r of
// lane l is Z(x = 4 r + l/16, y = l%16), which is input slot s = r of the transposed vector
// Z^T(x' = l%16, y' = l/16 + 4 s).  (E(x)E) acts alike on both lattice directions, so the next factor
// applied to Z^T gives the transpose of its product with Z: each factor takes the previous one's
// accumulators as its input with no data movement.  Only the site of a slot alternates between the
// two forms, and the diagonal scalings follow it:
//   normal     (V):   slot (l, s) is site l + 64 s
//   transposed (V^T): slot (l, s) is site l/16 + 4 s + 16 (l%16)
// The result is stored directly in the normal form and through a padded LDS transpose in the
// transposed one.  Per element each factor does what one kron_left16_kernel launch does
// (inner scaling, the two 16 x 16 x 16 products, then (z * cv) * row scaling); the sums of the
// products run over the lattice directions in the other order in the transposed form, so the
// result equals the launch-per-factor one to rounding.  Each wave carries 4 columns of X.
__global__ __launch_bounds__(256) void kron_chain16_kernel(KronChainArgs a)
{
    __shared__ double zt[4][16][17];
    const int b = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, q = lane >> 4;
    double ef[4], ef2[4];
    load_efrag<double>(a.E, lane, ef, ef2);
    const double* X = a.X + (long long)b * a.sX;
    double* C = a.C + (long long)b * a.sC;
    const int8_t* F = a.fields + (long long)b * a.fstride;
    const double dp = a.table[2 * b], dm = a.table[2 * b + 1];
    const int cbase = blockIdx.x * 16 + wave * 4;
    double v[4][4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
        for (int s = 0; s < 4; ++s) v[cc][s] = X[(long long)(cbase + cc) * 256 + lane + 64 * s];
    const int n = a.n;
    bool tr = false;   // form of the current vectors
    for (int k = 0; k < n; ++k) {
        const int8_t* f = F + (long long)a.ts[k] * 256;
        if (a.adjoint) {   // B^T = (E(x)E) diag(d): scale the input sites
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int site = tr ? (q + 4 * s + 16 * l16) : (lane + 64 * s);
                const double d = f[site] > 0 ? dp : dm;
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) v[cc][s] = v[cc][s] * d;
            }
        }
        d4_t z[4];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) kron16_vec<double>(ef, ef2, v[cc], z[cc]);
        tr = !tr;
        double dout[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int site = tr ? (q + 4 * r + 16 * l16) : (lane + 64 * r);
            dout[r] = a.adjoint ? 1.0 : (f[site] > 0 ? dp : dm);
        }
        const bool last = k == n - 1;
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
            const double cs = (last && a.cold.mode != DIAG_NONE) ? diag_value(a.cold, b, cbase + cc) : 1.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) v[cc][r] = (z[cc][r] * cs) * dout[r];
        }
    }
    // the squared norm of a column held at sites lane + 64 s: summed over s in order, then over the
    // lanes by wave_sum_dpp, as udv_rank_kernel sums a column it reads
    auto column_norm = [&](int col, const double (&zv)[4]) {
        double sq = 0.0;
#pragma unroll
        for (int s = 0; s < 4; ++s) sq = fma(zv[s], zv[s], sq);
        sq = wave_sum_dpp(sq);
        if (lane == 0) a.norms[(long long)b * 256 + col] = sq;
    };
    if (!tr) {
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
#pragma unroll
            for (int r = 0; r < 4; ++r) C[(long long)(cbase + cc) * 256 + lane + 64 * r] = v[cc][r];
            if (a.norms) column_norm(cbase + cc, v[cc]);
        }
    } else {
        // slot (l, r) holds site (4 r + l/16) + 16 (l%16): back to consecutive sites through LDS
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) {
#pragma unroll
            for (int r = 0; r < 4; ++r) zt[wave][l16][4 * r + q] = v[cc][r];
            wave_lds_sync();
            double zv[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) zv[s] = zt[wave][q + 4 * s][l16];
            wave_lds_sync();   // read before the next column overwrites the buffer
#pragma unroll
            for (int s = 0; s < 4; ++s) C[(long long)(cbase + cc) * 256 + lane + 64 * s] = zv[s];
            if (a.norms) column_norm(cbase + cc, zv);
        }
    }
}

hipError_t launch_kron_chain16(const KronChainArgs& a, hipStream_t s)
{
    if (a.n < 1 || a.n > KCHAIN_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kron_chain16_kernel, dim3(16, a.batch), dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_kron_left(const KronArgs<T>& a, hipStream_t s)
{
    if (a.nl > 16) return hipErrorInvalidValue;
    dim3 grid((a.N + 15) / 16, a.batch);
    if (a.nl == 16)
        hipLaunchKernelGGL(kron_left16_kernel<T>, grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(kron_left_kernel<T>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}
template <typename T>
hipError_t launch_kron_right(const KronArgs<T>& a, hipStream_t s)
{
    if (a.nl > 16) return hipErrorInvalidValue;
    dim3 grid((a.N + 15) / 16, a.batch);
    if (a.nl == 16)
        hipLaunchKernelGGL(kron_right16_kernel<T>, dim3(256 / KR16_ROWS, a.batch), dim3(512), 0, s, a);
    else
        hipLaunchKernelGGL(kron_right_kernel<T>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}
// exp(x A1) for the periodic ring of nl sites (A1[i][i+1] += 1, A1[i+1][i] += 1, so nl = 2 gives
// 2 and nl = 1 gives 2 on the diagonal), padded with zeros to a P x P row-major block.  The square
// lattice's exp(x A) is its Kronecker square.  Host code; a ring of more than P sites leaves out zero.
void ring_expk_padded(int nl, double x, int P, double* out)
{
    for (int e = 0; e < P * P; ++e) out[e] = 0.0;
    if (nl > P) return;
    if (x == 0.0) {   // exactly the identity (the cosine sums below round to ~1e-17)
        for (int i = 0; i < nl; ++i) out[i * P + i] = 1.0;
        return;
    }
    for (int i = 0; i < nl; ++i)
        for (int j = 0; j < nl; ++j) {
            double s = 0.0;
            for (int m = 0; m < nl; ++m) {
                const double k = 2.0 * M_PI * m / nl;
                s += std::cos(k * (i - j)) * std::exp(2.0 * x * std::cos(k));
            }
            out[i * P + j] = s / nl;
        }
}
void ring_expk16(int nl, double x, double* out) { ring_expk_padded(nl, x, 16, out); }

// Left product on lattices of 17..32 sites per side.  With E and every lattice vector V padded to
// 32 x 32 and split into 16 x 16 blocks, (E(x)E) v = E V E^T is computed as P = E V (P^T blocks
// accumulated from V^T E^T) and Z = P E^T, 32 + 32 MFMAs (16 x 16 x 4) per vector; as in the 16 x 16
// kernels, each block of P^T is already the A-fragment of the second step.  One wave per column of X:
// the column is read coalesced into the wave's LDS tile t[y][x] (row stride 33), the fragments are read
// from there, and Z goes back through the same tile for a coalesced store.  The diagonal scalings are
// applied at the load (inner) and the store (rows, column).
__global__ __launch_bounds__(256) void kron_left32_kernel(KronArgs<double> a)
{
    __shared__ double tl[4][32 * 33];
    typedef Mfma<double>::acc_t acc_t;
    const int b = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nl = a.nl, N = a.N, l16 = lane & 15, lq = lane >> 4;
    const int c = blockIdx.x * 4 + wave;
    if (c >= N) return;   // whole waves; the kernel has no workgroup barrier
    double* t = tl[wave];
    // B-fragments of E: e[p][q][s] holds E(16 p + l16, 16 q + lq + 4 s), i.e. B[k][col] of block (p, q)
    double e[2][2][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int s = 0; s < 4; ++s) e[p][q][s] = a.E[(16 * p + l16) * 32 + 16 * q + lq + 4 * s];
    for (int i = lane; i < 32 * 33; i += 64) t[i] = 0.0;
    wave_lds_sync();
    const double* X = a.X + (long long)b * a.sX + (long long)c * N;
    for (int site = lane; site < N; site += 64) {
        double v = X[site];
        if (a.innd.mode != DIAG_NONE) v *= diag_value(a.innd, b, site);
        t[(site / nl) * 33 + site % nl] = v;
    }
    wave_lds_sync();
    // step 1: acc1[bp][p] = block (y' in bp, x in p) of P^T; A(y', x') = V(x', y') = t[y'][x']
    acc_t acc1[2][2];
#pragma unroll
    for (int bp = 0; bp < 2; ++bp)
#pragma unroll
        for (int p = 0; p < 2; ++p) acc1[bp][p] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int bp = 0; bp < 2; ++bp)
#pragma unroll
        for (int cp = 0; cp < 2; ++cp)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const double vf = t[(16 * bp + l16) * 33 + 16 * cp + lq + 4 * s];
#pragma unroll
                for (int p = 0; p < 2; ++p) acc1[bp][p] = Mfma<double>::mma(vf, e[p][cp][s], acc1[bp][p]);
            }
    // step 2: acc2[p][q] = block (x in p, y in q) of Z; item s of acc1[bp][p] is A(x, y' = 16 bp + 4 s + lq)
    acc_t acc2[2][2];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            acc_t z = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int bp = 0; bp < 2; ++bp)
#pragma unroll
                for (int s = 0; s < 4; ++s) z = Mfma<double>::mma(acc1[bp][p][s], e[q][bp][s], z);
            acc2[p][q] = z;
        }
    wave_lds_sync();   // every fragment read precedes the writes of Z into the tile
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) t[(16 * q + l16) * 33 + 16 * p + 4 * r + lq] = acc2[p][q][r];
    wave_lds_sync();
    double* C = a.C + (long long)b * a.sC + (long long)c * N;
    const double cs = (a.cold.mode != DIAG_NONE) ? diag_value(a.cold, b, c) : 1.0;
    for (int site = lane; site < N; site += 64) {
        double v = t[(site / nl) * 33 + site % nl] * cs;
        if (a.rowd.mode != DIAG_NONE) v *= diag_value(a.rowd, b, site);
        C[site] = v;
    }
}

__global__ __launch_bounds__(256) void transpose_kernel(const double* __restrict__ Xs, long long sX,
                                                        double* __restrict__ Ys, long long sY, int N)
{
    __shared__ double tt[32][33];
    const int b = blockIdx.y, nt = (N + 31) / 32;
    const int i0 = (blockIdx.x % nt) * 32, j0 = (blockIdx.x / nt) * 32;
    const double* X = Xs + (long long)b * sX;
    double* Y = Ys + (long long)b * sY;
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int ii = e & 31, jj = e >> 5;   // X(i0 + ii, j0 + jj): consecutive ii contiguous
        tt[jj][ii] = (i0 + ii < N && j0 + jj < N) ? X[(i0 + ii) + (long long)(j0 + jj) * N] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int jj = e & 31, ii = e >> 5;   // Y(j0 + jj, i0 + ii) = X(i0 + ii, j0 + jj)
        if (i0 + ii < N && j0 + jj < N) Y[(j0 + jj) + (long long)(i0 + ii) * N] = tt[jj][ii];
    }
}

hipError_t launch_transpose(const double* X, long long sX, double* Y, long long sY, int N, int batch, hipStream_t s)
{
    const int nt = (N + 31) / 32;
    hipLaunchKernelGGL(transpose_kernel, dim3(nt * nt, batch), dim3(256), 0, s, X, sX, Y, sY, N);
    return hipGetLastError();
}

hipError_t launch_kron32_left(const KronArgs<double>& a, hipStream_t s)
{
    if (a.nl < 1 || a.nl > 32 || a.N != a.nl * a.nl) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kron_left32_kernel, dim3((a.N + 3) / 4, a.batch), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_kron32_right(const KronArgs<double>& a, double* S1, double* S2, hipStream_t s)
{
    hipError_t e = launch_transpose(a.X, a.sX, S1, a.sX, a.N, a.batch, s);
    if (e != hipSuccess) return e;
    KronArgs<double> t = a;
    t.X = S1; t.C = S2; t.sC = a.sX;
    t.rowd = a.cold;   // C^T = diag(cv) (E(x)E) diag(s) X^T diag(r)
    t.cold = a.rowd;
    if ((e = launch_kron32_left(t, s)) != hipSuccess) return e;
    return launch_transpose(S2, a.sX, a.C, a.sC, a.N, a.batch, s);
}

template hipError_t launch_kron_left<double>(const KronArgs<double>&, hipStream_t);
template hipError_t launch_kron_right<double>(const KronArgs<double>&, hipStream_t);
template hipError_t launch_kron_left<float>(const KronArgs<float>&, hipStream_t);
template hipError_t launch_kron_right<float>(const KronArgs<float>&, hipStream_t);

}  // namespace dqmc

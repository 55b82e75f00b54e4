// dqmc_kernels.hip -- gfx950 kernels of the DQMC sweep engine.
//
//   gemm_kernel ........ batched GEMM on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32)
//                        with fused diagonal scalings: the slice wraps G <- B G B^{-1}, the
//                        B-chain products and the UDV/recompose products.
//   metropolis_kernel .. one workgroup per chain walks the N sites of a slice: FP64 ratio
//                        from G_ii, minstd_rand0 uniforms by jump-ahead, delayed rank-k
//                        Sherman-Morrison updates flushed into G inside the workgroup.
//   qr_panel4_kernel ... blocked Householder QR of a row-major N x N block (panel in registers,
//   qr_update_kernel     trailing update on MFMA), applying Q^T to right-hand-side columns.
//   trsm_diag_kernel ... X = R^{-1} Z, 32-row diagonal blocks; the updates above run on MFMA.
//   small helpers ...... UDV pre-pivot / finish, system builders, error reductions, copies.
#include "dqmc_kernels.h"

#include <float.h>

#include <cmath>
#include <utility>

namespace dqmc {

// ---------------------------------------------------------------------------------------------
//  MFMA wrappers.  16x16x4 operand layout (64 lanes), measured on gfx950:
//    A fragment: lane l holds A[l%16][l/16]      B fragment: lane l holds B[l/16][l%16]
//    C fragment, f32: lane l, item r holds C[4*(l/16) + r][l%16]
//    C fragment, f64: lane l, item r holds C[4*r + l/16][l%16]
// ---------------------------------------------------------------------------------------------
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4_t acc_t;
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    typedef f4_t acc_t;
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// Chain-major block order for batched launches with `per` tiles per chain on a 1-D grid.  The
// dispatcher deals consecutive workgroups round-robin over the 8 XCDs, each with its own L2;
// when the chain count is a multiple of 8, workgroup L is given chain 8 (L/8 / per) + L%8 and
// tile (L/8) % per, so every tile of a chain runs on one XCD and the operands the tiles share
// are fetched from HBM into one L2 only.  Otherwise the plain order chain L/per, tile L%per.
__device__ __forceinline__ void chain_tile(int per, int batch, int& b, int& t)
{
    const int L = blockIdx.x;
    if ((batch & 7) == 0) {
        const int s = L >> 3;
        b = (s / per) * 8 + (L & 7);
        t = s % per;
    } else {
        b = L / per;
        t = L % per;
    }
}

__device__ __forceinline__ double diag_value(const DiagArg& d, int b, int i)
{
    switch (d.mode) {
        case DIAG_VEC: return d.v[b * d.vstride + i];
        case DIAG_INV: return 1.0 / d.v[b * d.vstride + i];
        case DIAG_INVMAX1: return 1.0 / fmax(1.0, d.v[b * d.vstride + i]);
        case DIAG_MIN1: return fmin(1.0, d.v[b * d.vstride + i]);
        case DIAG_FIELD: return d.f[b * d.fstride + i] > 0 ? d.v[b * d.vstride] : d.v[b * d.vstride + 1];
        default: return 1.0;
    }
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// LDS writes of some lanes of a wave become visible to the later LDS reads of the other lanes
// of the same wave (no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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

// E fragments B(k, col) = E[col][k], col = l%16.  Step 1 contracts k = l/16 + 4s.  Step 2
// contracts over the rows of step 1's accumulator, item s of lane l: k = Mfma<T>::row(l, s),
// i.e. 4s + l/16 for f64 and 4(l/16) + s for f32.
template <typename T>
__device__ __forceinline__ void load_efrag(const T* E, int lane, T (&ef)[4], T (&ef2)[4])
{
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        ef[s] = E[(lane & 15) * 16 + (lane >> 4) + 4 * s];
        ef2[s] = E[(lane & 15) * 16 + Mfma<T>::row(lane, s)];
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

// 16 x 16 lattice (N = 256): a lattice vector read as 64 lanes x 4 consecutive sites,
// site = l + 64 s, is exactly the A-fragment V(x = l%16, y = 4 s + l/16) of step 1, so
//   step 1  Y = V E^T   (A = V from a coalesced load, B(y, j) = E(j, y))
//   step 2  Z = E Y     (A(i, x) = E(i, x), B = step 1's accumulator items)
// and Z = E V E^T leaves the accumulator as Z(row(l, r), l%16), which a padded LDS transpose turns
// back into consecutive sites for a coalesced store.
template <typename T>
__device__ __forceinline__ void kron16_vec(const T (&ef)[4], const T (&ef2)[4], const T (&v)[4],
                                           typename Mfma<T>::acc_t& z)
{
    typedef typename Mfma<T>::acc_t acc_t;
    acc_t y = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; ++s) y = Mfma<T>::mma(v[s], ef[s], y);
    z = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 4; ++s) z = Mfma<T>::mma(ef2[s], y[s], z);
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
// 2 and nl = 1 gives 2 on the diagonal), padded to a 16 x 16 row-major block.  The square
// lattice's exp(x A) is its Kronecker square.  Host code.
void ring_expk16(int nl, double x, double* out)
{
    for (int e = 0; e < 256; ++e) out[e] = 0.0;
    if (x == 0.0) {   // exactly the identity (the cosine sums below round to ~1e-17)
        for (int i = 0; i < nl; ++i) out[i * 16 + i] = 1.0;
        return;
    }
    for (int i = 0; i < nl; ++i)
        for (int j = 0; j < nl; ++j) {
            double s = 0.0;
            for (int m = 0; m < nl; ++m) {
                const double k = 2.0 * M_PI * m / nl;
                s += std::cos(k * (i - j)) * std::exp(2.0 * x * std::cos(k));
            }
            out[i * 16 + j] = s / nl;
        }
}

template hipError_t launch_kron_left<double>(const KronArgs<double>&, hipStream_t);
template hipError_t launch_kron_right<double>(const KronArgs<double>&, hipStream_t);
template hipError_t launch_kron_left<float>(const KronArgs<float>&, hipStream_t);
template hipError_t launch_kron_right<float>(const KronArgs<float>&, hipStream_t);

// ---------------------------------------------------------------------------------------------
//  Metropolis with delayed updates.
//
//  For site i of slice t (sites in order 0..N-1, src/dqmc_core.cpp:106):
//    delta = e^{-2 alpha s} - 1,  r = (1 + (1 - G_ii) delta)^2 / (1 + delta)   (hubbard.cpp:123-130)
//    accept iff u < min(1, |r|), u = generate_canonical<double,53> of the chain's stream.
//  An accepted flip changes G by gamma * a b^T with gamma = delta / (1 + (1-G_ii) delta),
//  a = G(:,i), b = G(i,:)^T - e_i (hubbard.cpp:132-139).  The k pending rank-1 terms live
//  in LDS (X = gamma*a columns, Y = b rows); the current G is G0 + X Y, and G0 absorbs X Y
//  once k = kmax or at the end of the slice.  With kmax = 1 this is the reference's
//  immediate rank-1 update.
// ---------------------------------------------------------------------------------------------
#define MINSTD_M 2147483647ULL

__device__ __forceinline__ double canonical_from_draws(uint64_t x1, uint64_t x2)
{
    // generate_canonical<double,53>(minstd_rand0): sum = (x1-1) + (x2-1) * R (one fused
    // multiply-add), divided by fl(R^2) = 4611686009837453312, clamped below 1.
    const double R = 2147483646.0;
    const double R2 = 4611686009837453312.0;
    const double sum = __fma_rn((double)(x2 - 1ULL), R, (double)(x1 - 1ULL));
    double u = __ddiv_rn(sum, R2);
    if (u >= 1.0) u = 0x1.fffffffffffffp-1;
    return u;
}

__device__ __forceinline__ double accept_prob(double gii, double delta, double* xout)
{
#pragma clang fp contract(off)
    const double x = 1.0 + (1.0 - gii) * delta;
    const double r = x * x / (1.0 + delta);
    *xout = x;
    return fmin(1.0, fabs(r));
}

#define MB_MAX 32   // largest site block (delay) of the submatrix update
#define MB_LD 33    // row stride of XS / YS; column MB_MAX holds 0.0

// LDS layout of one Metropolis workgroup
template <typename T>
struct MetroLds {
    T* Xl;         // [kb][ys] X^T of the pending rank-a update (rows m; G row j of row m at j ^ yswz(m))
    T* Yl;         // [kb][ys] Y (rows m; column j of row m at j ^ yswz(m))
    int ys;        // row stride of Yl: N rounded up to 32
    double* colq;  // [32] column q of the current G(S,S), broadcast inside the site-loop wave
    double* rowq;  // [32] row q of the current G(S,S)
    double* XS;    // [MB_MAX][MB_LD]  gamma_m * G(S, i_m) of accepted flip m (restricted to the block)
    double* YS;    // [kb][kb]  G(i_m, S) - e_{i_m} (restricted to the block)
    double* unif;  // [N]
    double* gam;   // [kb]
    int* sl;       // [N] fields of the slice
    int* ql;       // [kb] block-local index of accepted flip m
    int* nacc;     // [2] accepted flips of site block j at nacc[j & 1]
    uint32_t* mask;
};

template <typename T>
__device__ MetroLds<T> metro_lds(unsigned char* smem, int N, int kb)
{
    MetroLds<T> L;
    const int kb4 = ((kb + 3) / 4) * 4;
    L.ys = ((N + 31) / 32) * 32;
    L.Xl = reinterpret_cast<T*>(smem);
    L.Yl = L.Xl + (size_t)kb4 * L.ys;
    L.colq = reinterpret_cast<double*>(L.Yl + (size_t)kb4 * L.ys);
    L.rowq = L.colq + 32;
    L.XS = L.rowq + 32;
    L.YS = L.XS + MB_MAX * MB_LD;
    L.unif = L.YS + MB_MAX * MB_LD;
    L.gam = L.unif + N;
    L.sl = reinterpret_cast<int*>(L.gam + kb);
    L.ql = L.sl + N;
    L.nacc = L.ql + kb;
    L.mask = reinterpret_cast<uint32_t*>(L.nacc + 2);
    return L;
}

size_t metropolis_lds_bytes(int N, int kb, size_t esz)
{
    const int kb4 = ((kb + 3) / 4) * 4, ys = ((N + 31) / 32) * 32;
    size_t b = (size_t)kb4 * 2 * ys * esz;
    b = (b + 7) / 8 * 8;
    b += (size_t)8 * (64 + 2 * MB_MAX * MB_LD + N + kb);
    b += (size_t)4 * (N + kb + 2) + (size_t)4 * ((N + 31) / 32) + 64;
    return b;
}

// Slot of row c of Xl and Yl: element (c, j) sits at c * ys + (j ^ yswz(c)).  The map stays inside
// the aligned 32-column block of j, so each column's slots belong to that column alone.  Two rows
// c, c + 1 of the same pair land in opposite halves of the 64 LDS banks, so the flush's 16-element
// reads of two rows at once, and the 16-lane x 2-row writes that stage rows, are free of bank
// conflicts.
__device__ __forceinline__ int yswz(int c) { return (c & ~1) ^ ((c & 1) << 4); }

// Tile sets of the flush, in 16 x 16 tiles (tile column ct, tile row rt; nt per side).  Tile rows and
// columns tr0..tr1 hold the next site block; its diagonal block (the G(S, S) its site loop reads) is
// FLUSH_DIAG, everything else FLUSH_OFFDIAG.  Consecutive tiles of a set share their tile column.
enum FlushTiles : int { FLUSH_ALL = 0, FLUSH_DIAG = 1, FLUSH_OFFDIAG = 2 };
__device__ __forceinline__ int flush_tile_count(int mode, int nt, int ncr)
{
    if (mode == FLUSH_ALL) return nt * nt;
    if (mode == FLUSH_DIAG) return ncr * ncr;
    return nt * nt - ncr * ncr;
}
__device__ __forceinline__ void flush_tile(int mode, int u, int nt, int tr0, int ncr, int& ct, int& rt)
{
    if (mode == FLUSH_ALL) {
        ct = u / nt;
        rt = u % nt;
    } else if (mode == FLUSH_DIAG) {
        ct = tr0 + u / ncr;
        rt = tr0 + u % ncr;
    } else if (u < (nt - ncr) * nt) {   // tile columns outside the block: every tile row
        ct = u / nt;
        rt = u % nt;
        if (ct >= tr0) ct += ncr;
    } else {                            // tile columns of the block: tile rows outside it
        const int v = u - (nt - ncr) * nt, o = nt - ncr;
        ct = tr0 + v / o;
        rt = v % o;
        if (rt >= tr0) rt += ncr;
    }
}

// G(ri, cj) += sum_m X[ri][m] Y[m][cj] over one tile set of the N x N matrix, on MFMA, by the nw
// waves w0 .. w0 + nw - 1 of the workgroup.  The tile is computed transposed (MFMA rows = G columns) so that each
// lane group touches contiguous column segments of the column-major G.  Each wave keeps FT tiles
// in flight (all loads of a group are issued before its first MFMA) so that the global-memory
// latency overlaps; FT is 8, or 4 under the 128-register budget of a 1024-thread workgroup.
template <typename T, int FT>
__device__ __forceinline__ void flush_mfma(T* __restrict__ G, int N, const T* __restrict__ Xl, const T* __restrict__ Yl,
                                           int ys, int a4, int mode, int tr0, int tr1, int w0, int nw)
{
    typedef typename Mfma<T>::acc_t acc_t;
    const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) - w0, l16 = lane & 15, lq = lane >> 4;
    const int nt = (N + 15) / 16, ncr = (mode == FLUSH_ALL) ? 0 : tr1 - tr0 + 1;
    const int ntt = flush_tile_count(mode, nt, ncr);
    for (int t0 = wave * FT; t0 < ntt; t0 += nw * FT) {
        acc_t acc[FT];
        int c0s[FT], r0s[FT];
        bool ok[FT];
#pragma unroll
        for (int u = 0; u < FT; ++u) {
            int ct = 0, rt = 0;
            ok[u] = t0 + u < ntt;
            if (ok[u]) flush_tile(mode, t0 + u, nt, tr0, ncr, ct, rt);
            c0s[u] = ok[u] ? ct * 16 : c0s[0];   // an empty slot repeats tile 0 and stores nothing
            r0s[u] = ok[u] ? rt * 16 : r0s[0];
        }
#pragma unroll
        for (int u = 0; u < FT; ++u) {
            const int gr = r0s[u] + l16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gc = c0s[u] + Mfma<T>::row(lane, r);
                acc[u][r] = (gr < N && gc < N) ? G[gr + (long long)gc * N] : T(0);
            }
        }
        // Operands: A(row = G column ca, k) = Yl[k][ca], B(k, col = G row gr) = Xl[k][gr], both swizzled.
        // Lanes past N read slots inside the padded rows; they only feed outputs that are not stored.
        for (int k0 = 0; k0 < a4; k0 += 4) {
            const int kk = k0 + lq, sw = yswz(kk);
            T av[FT], bv[FT];
#pragma unroll
            for (int u = 0; u < FT; ++u) {
                av[u] = Yl[kk * ys + ((c0s[u] + l16) ^ sw)];
                bv[u] = Xl[kk * ys + ((r0s[u] + l16) ^ sw)];
            }
#pragma unroll
            for (int u = 0; u < FT; ++u) acc[u] = Mfma<T>::mma(av[u], bv[u], acc[u]);
        }
#pragma unroll
        for (int u = 0; u < FT; ++u) {
            const int gr = r0s[u] + l16;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gc = c0s[u] + Mfma<T>::row(lane, r);
                if (ok[u] && gr < N && gc < N) G[gr + (long long)gc * N] = acc[u][r];
            }
        }
    }
}


// Accept/reject sequence of one site block S = b0 .. b0+bs-1 (bs <= 32) by one wave, with the
// reference's rank-1 update applied to G(S,S) after every accepted flip.  Lane l holds row
// r = l % 32 and columns 16 (l / 32) .. 16 (l / 32) + 15 of G(S,S) in registers, plus its own
// diagonal entry G(r, r) (kept by the same operations, so bit-identical to the block entry), and
// the acceptance probability and gamma of site r computed from it.  They change only when a flip
// is accepted, so every lane recomputes them then (in parallel), and a rejected site costs two
// lane reads and a compare.  An accepted flip q publishes column q and row q through LDS; every
// lane then applies G(r, c) += (gamma G(r, q)) (G(q, c) - delta_qc).  Records XS, YS, gam, ql and
// the flip count (in nacc[(b0 / kb) & 1]) for the rebuild, and flips the fields.
template <typename T>
__device__ __forceinline__ void metro_site_block(const MetroLds<T>& L, const T* __restrict__ G, int8_t* f, int N,
                                                 int kb, int b0, int bs, double dpos, double dneg, uint32_t* health)
{
    const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
    double g[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const int c = 16 * h + j;
        g[j] = (r < bs && c < bs) ? (double)G[(b0 + r) + (long long)(b0 + c) * N] : 0.0;
    }
    double dg = (r < bs) ? (double)G[(b0 + r) + (long long)(b0 + r) * N] : 0.0;
    const double uq = (r < bs) ? L.unif[b0 + r] : 0.0;
    const int sr = (r < bs) ? L.sl[b0 + r] : 1;
    const double dr = sr > 0 ? dpos : dneg;
    double xr, pr = accept_prob(dg, dr, &xr), gr = dr / xr;
    // A non-finite G_ii gives a non-finite ratio (the reference would take min(1, NaN) as 1);
    // within the block G(S,S) stays finite when it starts finite, so the block start is checked.
    if (__any(r < bs && !isfinite(xr)) && lane == 0) atomicOr(health, HEALTH_RATIO);
    int na = 0;
    for (int qh = 0; qh < 2; ++qh) {
        if (16 * qh >= bs) break;
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const int q = 16 * qh + jj;
            if (q >= bs) continue;
            if (readlane_f64(uq, q) < readlane_f64(pr, q)) {
                const double gam = readlane_f64(gr, q);
                if (h == qh) L.colq[r] = g[jj];
                if (r == q) {
#pragma unroll
                    for (int j = 0; j < 16; ++j) L.rowq[16 * h + j] = g[j];
                }
                wave_lds_sync();
                const double xs = gam * L.colq[r];
                const double ysd = L.rowq[r] - (r == q ? 1.0 : 0.0);
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int c = 16 * h + j;
                    const double ys = L.rowq[c] - (c == q ? 1.0 : 0.0);
                    g[j] += xs * ys;
                }
                dg += xs * ysd;
                pr = accept_prob(dg, dr, &xr);
                gr = dr / xr;
                if (lane < bs) {
                    L.XS[na * MB_LD + lane] = xs;
                    L.YS[na * MB_LD + lane] = ysd;
                }
                if (lane == 0) {
                    const int i = b0 + q;
                    const int si = L.sl[i];
                    L.gam[na] = gam;
                    L.ql[na] = q;
                    L.sl[i] = -si;
                    f[i] = (int8_t)(-si);
                    L.mask[i >> 5] |= 1u << (i & 31);
                }
                ++na;
                wave_lds_sync();
            }
        }
    }
    if (lane == 0) L.nacc[(b0 / kb) & 1] = na;
}

// acc + c[n] * x, where c[n] is lane n of the lane's 16-lane row (one v_fmac_f64_dpp with
// row_newbcast; every lane of the wave must be active).  The s_nop covers the two wait states a
// DPP read needs after a VALU write of its source, whatever the compiler placed before.
template <int n>
__device__ __forceinline__ double fmac_row_bcast(double acc, double c, double x);
#define DQMC_FMAC_ROW_BCAST(n)                                                                            \
    template <>                                                                                           \
    __device__ __forceinline__ double fmac_row_bcast<n>(double acc, double c, double x)                   \
    {                                                                                                     \
        asm volatile("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:" #n " row_mask:0xf bank_mask:0xf" \
                     : "+v"(acc)                                                                          \
                     : "v"(c), "v"(x));                                                                   \
        return acc;                                                                                       \
    }
DQMC_FMAC_ROW_BCAST(0)
DQMC_FMAC_ROW_BCAST(1)
DQMC_FMAC_ROW_BCAST(2)
DQMC_FMAC_ROW_BCAST(3)
DQMC_FMAC_ROW_BCAST(4)
DQMC_FMAC_ROW_BCAST(5)
DQMC_FMAC_ROW_BCAST(6)
DQMC_FMAC_ROW_BCAST(7)
DQMC_FMAC_ROW_BCAST(8)
DQMC_FMAC_ROW_BCAST(9)
DQMC_FMAC_ROW_BCAST(10)
DQMC_FMAC_ROW_BCAST(11)
DQMC_FMAC_ROW_BCAST(12)
DQMC_FMAC_ROW_BCAST(13)
DQMC_FMAC_ROW_BCAST(14)
DQMC_FMAC_ROW_BCAST(15)
#undef DQMC_FMAC_ROW_BCAST

// Step M of the right-looking form of the rebuild recursions: every later term m' > M (m' < na)
// takes its contribution from the finished term M, x[m'] = fma(c(m'), x[M], x[m']), with c(m')
// broadcast from lane m' % 16 of ca (m' < 16) or cb (m' >= 16).  Term m' receives its contributions in the order
// M = 0, 1, .., m'-1, the order of the left-looking sum x0[m'] + sum_{M<m'} x[M] c(m'), so both forms
// give the same bits.
template <int M, int P>
__device__ __forceinline__ void rebuild_term(double (&x)[MB_MAX], double ca, double cb)
{
    if constexpr (P > M) x[P] = fmac_row_bcast<(P & 15)>(x[P], P < 16 ? ca : cb, x[M]);
}
// Terms P0 .. P0+3 of step M; the whole group is skipped once P0 >= na (its terms stay zero).
template <int M, int P0>
__device__ __forceinline__ void rebuild_quad(double (&x)[MB_MAX], int na, double ca, double cb)
{
    if constexpr (P0 + 3 > M) {
        if (P0 < na) {
            rebuild_term<M, P0>(x, ca, cb);
            rebuild_term<M, P0 + 1>(x, ca, cb);
            rebuild_term<M, P0 + 2>(x, ca, cb);
            rebuild_term<M, P0 + 3>(x, ca, cb);
        }
    }
}
template <int M, int... Q>
__device__ __forceinline__ void rebuild_step(double (&x)[MB_MAX], int na, double ca, double cb, std::integer_sequence<int, Q...>)
{
    (rebuild_quad<M, 4 * Q>(x, na, ca, cb), ...);
}
template <bool SCALE, int M>
__device__ __forceinline__ void rebuild_one(double (&x)[MB_MAX], int na, const double* coef, int qa, int qb,
                                            const double* gam, double& ca, double& cb, double& g)
{
    if (SCALE) x[M] = g * x[M];
    double ta = 0.0, tb = 0.0, tg = 0.0;
    if (M + 1 < na) {   // the next step's operands, in flight during this step's products
        ta = coef[(M + 1) * MB_LD + qa];
        tb = coef[(M + 1) * MB_LD + qb];
        if (SCALE) tg = gam[M + 1];
    }
    rebuild_step<M>(x, na, ca, cb, std::make_integer_sequence<int, MB_MAX / 4>{});
    ca = ta;
    cb = tb;
    g = tg;
}
// The recursions of the delayed update, for one column per lane:
//   X:  x[m] = gam[m] (x0[m] + sum_{m'<m} x[m'] YS[m'][q_m])      (coef = YS, SCALE)
//   Y:  y[m] =         y0[m] + sum_{m'<m} XS[m'][q_m] y[m']       (coef = XS)
// coef is the [MB_MAX][MB_LD] LDS table of the block, row = accepted flip, column = block position,
// column MB_MAX = 0.0; qa / qb are q_{l % 16} and q_{16 + l % 16}, or MB_MAX past na, so that the
// coefficients of terms m >= na are zero and those terms stay zero.
template <bool SCALE, int... M>
__device__ __forceinline__ void rebuild_chain(double (&x)[MB_MAX], int na, const double* coef, int qa, int qb,
                                              const double* gam, std::integer_sequence<int, M...>)
{
    double ca = coef[qa], cb = coef[qb], g = SCALE ? gam[0] : 0.0;
    bool go = true;
    ((go = go && (M < na), go ? (rebuild_one<SCALE, M>(x, na, coef, qa, qb, gam, ca, cb, g), 0) : 0), ...);
}

// Submatrix Metropolis update of slice t.  Sites are visited in blocks S of kb consecutive
// sites.  Inside a block the accept/reject sequence runs on Gs = G(S,S) in LDS with the
// reference's rank-1 update.  After the block the accepted flips A are applied to the whole
// G as one rank-|A| product G <- G + X Y, where
//   X(:,m) = gamma_m (G0(:,i_m) + sum_{m'<m} X(:,m') Y(m',i_m))
//   Y(m,:) = G0(i_m,:) - e_{i_m} + sum_{m'<m} X(i_m,m') Y(m',:)
// are the sequential Sherman-Morrison vectors (src/hubbard.cpp:132-139) rebuilt from the
// columns/rows of G0 = G at block start; the block-restricted entries they need are the ones
// recorded in XS / YS while the block was processed.  With kb = 1 every accepted flip is
// applied immediately, as in the reference.
template <typename T, int MT>
__global__ __launch_bounds__(MT) __attribute__((amdgpu_waves_per_eu(MT / 256, MT / 256))) void metropolis_kernel(MetroArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, kb = a.kmax;
    const int b = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    MetroLds<T> L = metro_lds<T>(smem, N, kb);

    T* G = a.G + (long long)b * a.sG;
    int8_t* f = a.fields + ((long long)b * a.L + a.t) * N;
    const uint64_t x0 = a.rng[b];
    const int nw = (N + 31) / 32;
    for (int j = tid; j < N; j += nth) {
        L.sl[j] = f[j];
        const uint64_t x1 = (x0 * a.powtab[2 * j]) % MINSTD_M;
        const uint64_t x2 = (x0 * a.powtab[2 * j + 1]) % MINSTD_M;
        L.unif[j] = canonical_from_draws(x1, x2);
    }
    for (int w = tid; w < nw; w += nth) L.mask[w] = 0u;
    for (int w = tid; w < MB_MAX; w += nth) L.XS[w * MB_LD + MB_MAX] = L.YS[w * MB_LD + MB_MAX] = 0.0;

    const double dpos = a.delta[2 * b], dneg = a.delta[2 * b + 1];
    int nacc = 0;
    // Software pipeline over the site blocks: after block B's rebuild, wave 0 applies B's flips to the
    // diagonal block G(S', S') of block B+1 and runs B+1's site loop, which reads only that block,
    // while the other waves apply them to the rest of G (the rows and columns of B+1 included, which
    // B+1's rebuild reads after the next barrier).
    __syncthreads();
    uint32_t* health = a.health + b;
    if (tid < 64) metro_site_block(L, G, f, N, kb, 0, min(kb, N), dpos, dneg, health);
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, nwv = nth >> 6, l16 = lane & 15;
    for (int b0 = 0; b0 < N; b0 += kb) {
        const int bs = min(kb, N - b0);
        // The count of block b0 sits in its own parity slot: when na == 0 no barrier separates this
        // read from the site loop of the next block, which writes the other slot.
        const int na = L.nacc[(b0 / kb) & 1];
        nacc += na;
        const int b1 = b0 + kb, bs1 = (b1 < N) ? min(kb, N - b1) : 0;
        const bool next = b1 < N;
        const int tr0 = next ? b1 / 16 : 0, tr1 = next ? (b1 + bs1 - 1) / 16 : 0;
        const int a4 = ((na + 3) / 4) * 4;
        if (na > 0) {
            // Rebuild X (columns) and Y (rows) of the rank-na update from G0.  All waves first stage
            // the block's rows G0(S, :) in Yl (16 lanes per column, two rows per lane: contiguous
            // global reads, all in flight at once).  Then each wave takes whole 64-column tasks, X
            // or Y, and runs the recursion of its columns in registers with all 64 lanes active (the
            // coefficients travel by DPP broadcast, which needs every source lane).
            const int ys = L.ys, nct = (N + 63) / 64;
            const int qlane = (lane < na) ? L.ql[lane] : 0;   // lane m holds q_m
            const int qa = (l16 < na) ? L.ql[l16] : MB_MAX, qb = (16 + l16 < na) ? L.ql[16 + l16] : MB_MAX;
            {
                constexpr int SI = (16 * 256 + MT - 1) / MT;   // staging items per thread at N <= 256
                const int r = 2 * l16, r0 = min(r, bs - 1), r1 = min(r + 1, bs - 1);
                T v0[SI], v1[SI];
#pragma unroll
                for (int i = 0; i < SI; ++i) {
                    const int jj = min(((tid + i * nth) >> 4), N - 1);
                    v0[i] = G[(b0 + r0) + (long long)jj * N];
                    v1[i] = G[(b0 + r1) + (long long)jj * N];
                }
#pragma unroll
                for (int i = 0; i < SI; ++i) {
                    const int jj = (tid + i * nth) >> 4;
                    if (jj < N && r < bs) L.Yl[r * ys + (jj ^ yswz(r))] = v0[i];
                    if (jj < N && r + 1 < bs) L.Yl[(r + 1) * ys + (jj ^ yswz(r + 1))] = v1[i];
                }
                for (int w = tid + SI * nth; w < N * 16; w += nth) {   // workgroups narrower than MT
                    const int jj = w >> 4;
                    if (r < bs) L.Yl[r * ys + (jj ^ yswz(r))] = G[(b0 + r) + (long long)jj * N];
                    if (r + 1 < bs) L.Yl[(r + 1) * ys + (jj ^ yswz(r + 1))] = G[(b0 + r + 1) + (long long)jj * N];
                }
            }
            __syncthreads();
            for (int task = wave; task < 2 * nct; task += nwv) {
                const bool isx = task < nct;
                const int j = (isx ? task : task - nct) * 64 + lane;
                const bool jin = j < N;
                double xr[MB_MAX];
                if (isx) {
#pragma unroll
                    for (int m = 0; m < MB_MAX; ++m) {
                        const int q = __builtin_amdgcn_readlane(qlane, m);
                        xr[m] = (m < na && jin) ? (double)G[j + (long long)(b0 + q) * N] : 0.0;
                    }
                    rebuild_chain<true>(xr, na, L.YS, qa, qb, L.gam, std::make_integer_sequence<int, MB_MAX>{});
#pragma unroll
                    for (int m = 0; m < MB_MAX; ++m)
                        if (m < a4 && jin) L.Xl[m * ys + (j ^ yswz(m))] = T(xr[m]);
                } else {
#pragma unroll
                    for (int m = 0; m < MB_MAX; ++m) {
                        const int q = __builtin_amdgcn_readlane(qlane, m);
                        xr[m] = (m < na && jin) ? (double)L.Yl[q * ys + (j ^ yswz(q))] - (j == b0 + q ? 1.0 : 0.0) : 0.0;
                    }
                    rebuild_chain<false>(xr, na, L.XS, qa, qb, L.gam, std::make_integer_sequence<int, MB_MAX>{});
#pragma unroll
                    for (int m = 0; m < MB_MAX; ++m)
                        if (m < a4 && jin) L.Yl[m * ys + (j ^ yswz(m))] = T(xr[m]);
                }
            }
            __syncthreads();
        }
        if (next) {
            if (wave == 0) {
                if (na > 0) {
                    flush_mfma<T, (MT >= 1024 ? 4 : 8)>(G, N, L.Xl, L.Yl, L.ys, a4, FLUSH_DIAG, tr0, tr1, 0, 1);
                    // this wave's stores of G(S', S') complete before its site loop reads them
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                metro_site_block(L, G, f, N, kb, b1, bs1, dpos, dneg, health);
            } else if (na > 0) {
                flush_mfma<T, (MT >= 1024 ? 4 : 8)>(G, N, L.Xl, L.Yl, L.ys, a4, FLUSH_OFFDIAG, tr0, tr1, 1, nwv - 1);
            }
        } else if (na > 0) {
            flush_mfma<T, (MT >= 1024 ? 4 : 8)>(G, N, L.Xl, L.Yl, L.ys, a4, FLUSH_ALL, 0, 0, 0, nwv);
        }
        __syncthreads();
    }
    __syncthreads();
    uint32_t* gm = a.accmask + ((long long)b * a.L + a.t) * nw;
    for (int w = tid; w < nw; w += nth) gm[w] = L.mask[w];
    if (tid == 0) {
        a.rng[b] = (x0 * a.powtab[2 * N - 1]) % MINSTD_M;
        a.acc_count[b] += (unsigned long long)nacc;
    }
}



hipError_t prepare_kernels()
{
    const void* ks[] = {(const void*)metropolis_kernel<double, 256>, (const void*)metropolis_kernel<double, 512>,
                        (const void*)metropolis_kernel<double, 1024>, (const void*)metropolis_kernel<float, 256>,
                        (const void*)metropolis_kernel<float, 512>, (const void*)metropolis_kernel<float, 1024>};
    for (const void* k : ks) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename T>
hipError_t launch_metropolis(const MetroArgs<T>& a, int batch, int threads, hipStream_t s)
{
    if (a.kmax < 1 || a.kmax > MB_MAX) return hipErrorInvalidValue;
    const size_t shm = metropolis_lds_bytes(a.N, a.kmax, sizeof(T));
    if (shm > 160 * 1024) return hipErrorInvalidValue;
    if (a.N < 256) threads = ((a.N + 63) / 64) * 64 * 2 <= 1024 ? ((a.N + 63) / 64) * 64 * 2 : 1024;
    if (threads <= 256)
        hipLaunchKernelGGL((metropolis_kernel<T, 256>), dim3(batch), dim3(threads < 128 ? 128 : threads), shm, s, a);
    else if (threads <= 512)
        hipLaunchKernelGGL((metropolis_kernel<T, 512>), dim3(batch), dim3(threads), shm, s, a);
    else
        hipLaunchKernelGGL((metropolis_kernel<T, 1024>), dim3(batch), dim3(1024), shm, s, a);
    return hipGetLastError();
}
template hipError_t launch_metropolis<double>(const MetroArgs<double>&, int, int, hipStream_t);
template hipError_t launch_metropolis<float>(const MetroArgs<float>&, int, int, hipStream_t);

// ---------------------------------------------------------------------------------------------
//  Cross-lane reductions shared by the QR kernels.
// ---------------------------------------------------------------------------------------------
#define QR_NB 16      // panel width of the one-kernel QR (qr_v1_kernel)
#define QR_MAXN 256

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Register-only cross-lane moves of f64 values (DPP and the gfx950 permlane swaps), measured lane
// patterns: quad_perm [1,0,3,2] = lane ^ 1, quad_perm [2,3,0,1] = lane ^ 2, row_half_mirror = lane ^ 7
// within 8 lanes, row_mirror = lane ^ 15 within 16 lanes; permlane16_swap(v, v) returns the values
// of the even and the odd row of each row pair, permlane32_swap(v, v) those of the two halves.
#define DPP_XOR1 0xB1
#define DPP_XOR2 0x4E
#define DPP_HALF_MIRROR 0x141
#define DPP_MIRROR 0x140
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// v(lane) + v(lane ^ 16), added in the same order in both lanes
__device__ __forceinline__ double sum_xor16(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(u >> 32), (unsigned)(u >> 32), false, false);
    const double even = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    const double odd = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
    return even + odd;
}
// v(lane) + v(lane ^ 32), added in the same order in both lanes
__device__ __forceinline__ double sum_xor32(double v)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)u, (unsigned)u, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(u >> 32), (unsigned)(u >> 32), false, false);
    const double low = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    const double high = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
    return low + high;
}
// sum over the 16 lanes of a row, bit-identical in all of them
__device__ __forceinline__ double row_sum_dpp(double v)
{
    v += dpp_f64<DPP_XOR1>(v);
    v += dpp_f64<DPP_XOR2>(v);
    v += dpp_f64<DPP_HALF_MIRROR>(v);
    v += dpp_f64<DPP_MIRROR>(v);
    return v;
}
__device__ __forceinline__ double wave_sum_dpp(double v) { return sum_xor32(sum_xor16(row_sum_dpp(v))); }



// ---------------------------------------------------------------------------------------------
//  One-kernel Householder QR (DQMC_QR=v1): one 256-thread workgroup per chain factors panels of
//  QR_NB columns in LDS and updates the trailing columns thread-per-column through global memory.
// ---------------------------------------------------------------------------------------------


__global__ __launch_bounds__(256) void qr_v1_kernel(double* __restrict__ bufs, long long stride, int N, int W)
{
    __shared__ double P[QR_MAXN][QR_NB + 1];  // panel rows p0..N-1 (relative), reflectors below diag
    __shared__ double Tm[QR_NB][QR_NB];
    __shared__ double VtV[QR_NB][QR_NB];
    __shared__ double tau[QR_NB];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double* A = bufs + (long long)blockIdx.x * stride;

    for (int p0 = 0; p0 < N; p0 += QR_NB) {
        const int nb = min(QR_NB, N - p0);
        const int R = N - p0;  // panel rows
        // load panel
        for (int e = tid; e < R * nb; e += 256) {
            const int r = e / nb, c = e % nb;
            P[r][c] = A[(long long)(p0 + r) * W + p0 + c];
        }
        __syncthreads();
        // factor panel column by column
        for (int c = 0; c < nb; ++c) {
            if (wave == 0) {
                double s = 0.0;
                for (int r = c + 1 + lane; r < R; r += 64) s += P[r][c] * P[r][c];
                s = wave_sum(s);
                const double alpha = P[c][c];
                double tc = 0.0, beta = alpha, scal = 0.0;
                if (s > 0.0) {
                    beta = -copysign(sqrt(alpha * alpha + s), alpha);
                    tc = (beta - alpha) / beta;
                    scal = 1.0 / (alpha - beta);
                }
                for (int r = c + 1 + lane; r < R; r += 64) P[r][c] *= scal;
                if (lane == 0) { P[c][c] = beta; tau[c] = tc; }
            }
            __syncthreads();
            // apply H_c to panel columns c+1..nb-1: one wave per column
            for (int cc = c + 1 + wave; cc < nb; cc += 4) {
                double s = 0.0;
                for (int r = c + lane; r < R; r += 64) s += (r == c ? 1.0 : P[r][c]) * P[r][cc];
                s = wave_sum(s) * tau[c];
                for (int r = c + lane; r < R; r += 64) P[r][cc] -= s * (r == c ? 1.0 : P[r][c]);
            }
            __syncthreads();
        }
        // write back R rows of the panel (upper part) and zeros below the diagonal
        for (int e = tid; e < R * nb; e += 256) {
            const int r = e / nb, c = e % nb;
            A[(long long)(p0 + r) * W + p0 + c] = (r <= c) ? P[r][c] : 0.0;
        }
        // V^T V for the compact WY factor (V unit lower trapezoidal)
        for (int e = tid; e < nb * nb; e += 256) {
            const int i = e / nb, j = e % nb;
            double s = 0.0;
            if (i <= j) {
                // rows r >= j (v_i(r) nonzero for r >= i, v_j for r >= j)
                for (int r = j; r < R; ++r) {
                    const double vi = (r == i) ? 1.0 : P[r][i];
                    const double vj = (r == j) ? 1.0 : P[r][j];
                    s += vi * vj;
                }
            }
            VtV[i][j] = s;
        }
        __syncthreads();
        if (tid == 0) {
            // T(0:i,i) = -tau_i T(0:i,0:i) V(:,0:i)^T v_i ; T(i,i) = tau_i  (forward, columnwise)
            for (int i = 0; i < nb; ++i) {
                for (int r = 0; r < i; ++r) {
                    double s = 0.0;
                    for (int q = r; q < i; ++q) s += Tm[r][q] * VtV[q][i];
                    Tm[r][i] = -tau[i] * s;
                }
                Tm[i][i] = tau[i];
                for (int r = i + 1; r < nb; ++r) Tm[r][i] = 0.0;
            }
        }
        __syncthreads();
        // trailing update of columns p0+nb .. W-1: A <- A - V T^T V^T A
        for (int col = p0 + nb + tid; col < W; col += 256) {
            double w[QR_NB];
#pragma unroll
            for (int m = 0; m < QR_NB; ++m) w[m] = 0.0;
            double* Acol = A + (long long)p0 * W + col;
            for (int r = 0; r < R; ++r) {
                const double x = Acol[(long long)r * W];
#pragma unroll
                for (int m = 0; m < QR_NB; ++m) {
                    if (m < nb) {
                        const double v = (r == m) ? 1.0 : (r > m ? P[r][m] : 0.0);
                        w[m] += v * x;
                    }
                }
            }
            double wt[QR_NB];
#pragma unroll
            for (int m = 0; m < QR_NB; ++m) {
                double s = 0.0;
#pragma unroll
                for (int q = 0; q < QR_NB; ++q)
                    if (q <= m && m < nb) s += Tm[q][m] * w[q];
                wt[m] = s;
            }
            for (int r = 0; r < R; ++r) {
                double s = 0.0;
#pragma unroll
                for (int m = 0; m < QR_NB; ++m) {
                    if (m < nb) {
                        const double v = (r == m) ? 1.0 : (r > m ? P[r][m] : 0.0);
                        s += v * wt[m];
                    }
                }
                Acol[(long long)r * W] -= s;
            }
        }
        __syncthreads();
    }
}

hipError_t launch_qr_v1(double* buf, long long stride, int N, int W, int batch, hipStream_t s)
{
    if (N > QR_MAXN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(qr_v1_kernel, dim3(batch), dim3(256), 0, s, buf, stride, N, W);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
//  Blocked Householder QR in four launches per panel of QRB columns:
//    qr_panel4_kernel one 1024-thread workgroup per chain keeps the panel rows in registers,
//                     factors it column by column (Householder, LAPACK dlarfg convention) and
//                     forms the compact-WY factor T; writes R back, V and T out.
//    GEMM 1 (MFMA)    Wt = At^T V         (At = trailing columns of the panel rows)
//    GEMM 2 (MFMA)    Wt2 = Wt T
//    GEMM 3 (MFMA)    At^T -= Wt2 V^T,  i.e. At <- (I - V T^T V^T) At = Q_panel^T At
//  The trailing columns include the right-hand sides, so they receive Q^T rhs.
// ---------------------------------------------------------------------------------------------
#define QRB 32

// Panel factorization with four threads per panel row (1024 threads; thread t handles row
// t % 256 and the 8 panel columns 8 (t / 256) .. 8 (t / 256) + 7), so that each column step costs
// each thread O(8) work and every SIMD holds four waves.  Writes R back, V (as above) and the
// compact-WY factor T (column-major QRB x QRB per chain) with H_1..H_nb = I - V T V^T.
#define QCPT 8
__global__ __launch_bounds__(1024) void qr_panel4_kernel(double* __restrict__ bufs, long long stride, int N, int W,
                                                         int p0, int nb, double* __restrict__ Vs,
                                                         double* __restrict__ Ts, double* __restrict__ Vrows,
                                                         long long vr_stride)
{
    __shared__ double red[16][QCPT];        // per-wave partial dot products
    __shared__ double red1[16];             // per-wave partial norms
    __shared__ double xcol[256];            // column i of the panel (before its reflection)
    __shared__ double VtV[QRB][QRB + 1];    // VtV[k][i] = v_k . v_i, k < i
    __shared__ double tau_s[QRB];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int g = t >> 8, r = t & 255;
    const int b = blockIdx.x;
    const int R = N - p0;
    const bool live = r < R;
    double* A = bufs + (long long)b * stride;
    double* V = Vs + (long long)b * N * QRB;
    double* Tg = Ts + (long long)b * QRB * QRB;

    double x[QCPT];
#pragma unroll
    for (int m = 0; m < QCPT; ++m) {
        const int col = g * QCPT + m;
        x[m] = (live && col < nb) ? A[(long long)(p0 + r) * W + p0 + col] : 0.0;
    }

    for (int i = 0; i < nb; ++i) {
        const int go = i / QCPT, li = i % QCPT;
        double xi = 0.0;
#pragma unroll
        for (int m = 0; m < QCPT; ++m)
            if (m == li) xi = x[m];
        const bool owner = g == go;
        if (owner) xcol[r] = live ? xi : 0.0;
        double n2 = (owner && live && r > i) ? xi * xi : 0.0;
        n2 = wave_sum_dpp(n2);
        if (lane == 0) red1[wave] = n2;
        __syncthreads();
        double s2 = 0.0;
#pragma unroll
        for (int w = 0; w < 4; ++w) s2 += red1[go * 4 + w];
        const double alpha = xcol[i];
        double beta = alpha, tau = 0.0, scal = 0.0;
        if (s2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + s2), alpha);
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        const double vr = (r == i) ? 1.0 : ((live && r > i) ? xcol[r] * scal : 0.0);
        if (owner) {
            const double xnew = (r == i) ? beta : ((r > i) ? vr : xi);
#pragma unroll
            for (int m = 0; m < QCPT; ++m)
                if (m == li) x[m] = xnew;
        }
        if (t == 0) tau_s[i] = tau;
        // this thread's share of v_m . v_i (m < i) and v_i . x_m (m > i)
        double d[QCPT];
#pragma unroll
        for (int m = 0; m < QCPT; ++m) {
            const int col = g * QCPT + m;
            double vm = 0.0;
            if (col < i) vm = (r == col) ? 1.0 : (r > col ? x[m] : 0.0);
            else if (col > i && col < nb) vm = x[m];
            d[m] = vm * vr;
        }
        // recursive halving over the 64 lanes, register-only: 8 values -> 4 (pairs across lane bit 3),
        // 2 (bit 2), 1 (bit 1), then plain sums across bits 0, 4 and 5; lane owns index
        // 4 bit3 + 2 bit2 + bit1.
        {
            double w4[4], w2[2], w1;
            const bool h8 = lane & 8, h4 = lane & 4, h2 = lane & 2;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const double keep = h8 ? d[m + 4] : d[m], send = h8 ? d[m] : d[m + 4];
                w4[m] = keep + dpp_f64<DPP_MIRROR>(send);
            }
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const double keep = h4 ? w4[m + 2] : w4[m], send = h4 ? w4[m] : w4[m + 2];
                w2[m] = keep + dpp_f64<DPP_HALF_MIRROR>(send);
            }
            {
                const double keep = h2 ? w2[1] : w2[0], send = h2 ? w2[0] : w2[1];
                w1 = keep + dpp_f64<DPP_XOR2>(send);
            }
            w1 += dpp_f64<DPP_XOR1>(w1);
            w1 = sum_xor32(sum_xor16(w1));
            const int idx = (h8 ? 4 : 0) + (h4 ? 2 : 0) + (h2 ? 1 : 0);
            if ((lane & 0x31) == 0) red[wave][idx] = w1;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < QCPT; ++m) {
            const int col = g * QCPT + m;
            if (col > i && col < nb) {
                const double dm = red[g * 4][m] + red[g * 4 + 1][m] + red[g * 4 + 2][m] + red[g * 4 + 3][m];
                x[m] -= tau * vr * dm;
            }
        }
        if (t < i) {
            const int gk = t / QCPT, lk = t % QCPT;
            VtV[t][i] = red[gk * 4][lk] + red[gk * 4 + 1][lk] + red[gk * 4 + 2][lk] + red[gk * 4 + 3][lk];
        }
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int m = 0; m < QCPT; ++m) {
            const int col = g * QCPT + m;
            if (col < nb) {
                A[(long long)(p0 + r) * W + p0 + col] = (r <= col) ? x[m] : 0.0;
                V[r + (long long)col * N] = (r == col) ? 1.0 : (r > col ? x[m] : 0.0);
            }
        }
        if (Vrows) {
            // the same V row-major (row r at r * QRB), for row-fragment reads of the update
            double* vr = Vrows + (long long)b * vr_stride + (long long)r * QRB;
#pragma unroll
            for (int m = 0; m < QCPT; ++m) {
                const int col = g * QCPT + m;
                if (col < nb) vr[col] = (r == col) ? 1.0 : (r > col ? x[m] : 0.0);
            }
        }
    }
    // T(j, j) = tau_j, T(j, i) = -tau_i sum_{k=j}^{i-1} T(j, k) VtV(k, i): lane (j, k) = (t / 32, t % 32)
    // holds T(j, k); the 32 column steps each sum one row over its 32 lanes in registers.
    {
        const int j = t >> 5, k = t & 31;
        double tk = 0.0;
        for (int i = 0; i < nb; ++i) {
            double c = (k >= j && k < i) ? tk * VtV[k][i] : 0.0;
            c = sum_xor16(row_sum_dpp(c));
            const double tji = (i == j) ? tau_s[i] : ((i > j) ? -tau_s[i] * c : 0.0);
            if (k == i) tk = tji;
        }
        Tg[j + (long long)k * QRB] = tk;
    }
}

// Sum every one of the 32 values d[0..31] over the 64 lanes of a wave by recursive halving, with
// register-only exchanges: at each level a lane keeps the half of its values selected by one lane
// bit and adds the same half of the partner lane that differs in that bit.  For bits 5 and 4 one
// permlane32 / permlane16 swap of the pair (low half, high half) hands every lane its own kept value
// and the partner's in the two results, so their sum needs no select; bits 3 and 2 use the DPP row
// mirror and half-row mirror, bits 1 and 0 quad perms.  On exit lane l holds the wave sum of value
// 16 b5 + 8 b4 + 4 b3 + 2 b2 + b1 (b_k = bit k of l), identical in lanes l and l ^ 1.
template <bool ROW16>
__device__ __forceinline__ double swap_sum_f64(double lo_half, double hi_half)
{
    const unsigned long long a = (unsigned long long)__double_as_longlong(lo_half);
    const unsigned long long b = (unsigned long long)__double_as_longlong(hi_half);
    unsigned r0l, r1l, r0h, r1h;
    if (ROW16) {
        const auto l = __builtin_amdgcn_permlane16_swap((unsigned)a, (unsigned)b, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap((unsigned)(a >> 32), (unsigned)(b >> 32), false, false);
        r0l = l[0]; r1l = l[1]; r0h = h[0]; r1h = h[1];
    } else {
        const auto l = __builtin_amdgcn_permlane32_swap((unsigned)a, (unsigned)b, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap((unsigned)(a >> 32), (unsigned)(b >> 32), false, false);
        r0l = l[0]; r1l = l[1]; r0h = h[0]; r1h = h[1];
    }
    const double r0 = __longlong_as_double((long long)(((unsigned long long)r0h << 32) | r0l));
    const double r1 = __longlong_as_double((long long)(((unsigned long long)r1h << 32) | r1l));
    return r0 + r1;
}
__device__ __forceinline__ double wave_halve32(const double (&d)[32], int lane)
{
    double a16[16], a8[8], a4[4], a2[2];
    const bool h8 = lane & 8, h4 = lane & 4, h2 = lane & 2;
#pragma unroll
    for (int m = 0; m < 16; ++m) a16[m] = swap_sum_f64<false>(d[m], d[m + 16]);
#pragma unroll
    for (int m = 0; m < 8; ++m) a8[m] = swap_sum_f64<true>(a16[m], a16[m + 8]);
#pragma unroll
    for (int m = 0; m < 4; ++m)
        a4[m] = (h8 ? a8[m + 4] : a8[m]) + dpp_f64<DPP_MIRROR>(h8 ? a8[m] : a8[m + 4]);
#pragma unroll
    for (int m = 0; m < 2; ++m)
        a2[m] = (h4 ? a4[m + 2] : a4[m]) + dpp_f64<DPP_HALF_MIRROR>(h4 ? a4[m] : a4[m + 2]);
    double a1 = (h2 ? a2[1] : a2[0]) + dpp_f64<DPP_XOR2>(h2 ? a2[0] : a2[1]);
    return a1 + dpp_f64<DPP_XOR1>(a1);
}
__device__ __forceinline__ int halve32_index(int lane)
{
    return ((lane & 32) ? 16 : 0) + ((lane & 16) ? 8 : 0) + ((lane & 8) ? 4 : 0) + ((lane & 4) ? 2 : 0) +
           ((lane & 2) ? 1 : 0);
}

// Panel factorization, 256 threads (4 waves) per chain: thread r holds panel row r (rows p0 .. N-1,
// so N - p0 <= 256) with all QRB columns in registers.  Column i costs two 4-wave barriers: one
// after the partial norms (alpha = x(i,i) is published with them), one after the partial sums of
// the 32 products v . x_j (j > i) and v_k . v (k < i, the V^T V entries of T) that one recursive
// halving reduces per wave.  The column sums reach every lane through v_readlane (uniform values).
// Every LDS slot is written between the second barrier of one column and the first of the next,
// and read between the two barriers of the column that wrote it, so single buffers suffice.
// T (H_1..H_nb = I - V T V^T, column-major QRB x QRB) comes from the recursion
// T(j,i) = -tau_i sum_{k=j}^{i-1} T(j,k) (V^T V)(k,i), thread j forming row j on its own.
__global__ __launch_bounds__(256) void qr_panel_kernel(double* __restrict__ bufs, long long stride, int N, int W,
                                                       int p0, int nb, double* __restrict__ Vs,
                                                       double* __restrict__ Ts, double* __restrict__ Vrows,
                                                       long long vr_stride)
{
    __shared__ double nrm[4];
    __shared__ double alph;
    __shared__ double dot[4][QRB];
    __shared__ __align__(16) double VtVt[QRB][QRB];   // VtVt[i][k] = (V^T V)(k, i), k < i
    __shared__ double tau_s[QRB];
    const int r = threadIdx.x, lane = r & 63, wave = r >> 6;
    const int b = blockIdx.x;
    const int R = N - p0;
    const bool live = r < R;
    double* A = bufs + (long long)b * stride;
    double* V = Vs + (long long)b * N * QRB;
    double* Tg = Ts + (long long)b * QRB * QRB;

    double x[QRB];
#pragma unroll
    for (int m = 0; m < QRB; ++m) x[m] = (live && m < nb) ? A[(long long)(p0 + r) * W + p0 + m] : 0.0;
    if (r < QRB) tau_s[r] = 0.0;
    for (int e = r; e < QRB * QRB; e += 256) VtVt[e / QRB][e % QRB] = 0.0;

#pragma clang loop unroll(full)
    for (int i = 0; i < QRB; ++i) {
        if (i >= nb) continue;   // uniform
        const double xi = x[i];
        if (r == i) alph = xi;
        double n2 = (live && r > i) ? xi * xi : 0.0;
        n2 = wave_sum_dpp(n2);
        if (lane == 0) nrm[wave] = n2;
        __syncthreads();
        const double s2 = ((nrm[0] + nrm[1]) + nrm[2]) + nrm[3];
        const double alpha = alph;
        double beta = alpha, tau = 0.0, scal = 0.0;
        if (s2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + s2), alpha);
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        const double v = (r == i) ? 1.0 : ((live && r > i) ? xi * scal : 0.0);
        if (r == i) x[i] = beta;
        else if (r > i) x[i] = v;
        // m < i: v_m . v_i (V^T V (m, i)); m > i: v_i . x_m.  Rows r < i have v = 0, and rows r >= i
        // hold v_m(r) in x[m] for m < i; columns m >= nb stay zero, so no product needs a guard.
        double d[QRB];
#pragma unroll
        for (int m = 0; m < QRB; ++m) d[m] = (m == i) ? 0.0 : v * x[m];
        const double ws = wave_halve32(d, lane);
        if ((lane & 1) == 0) dot[wave][halve32_index(lane)] = ws;
        if (r == 0) tau_s[i] = tau;
        __syncthreads();
        const double wsum = (lane < QRB) ? ((dot[0][lane] + dot[1][lane]) + dot[2][lane]) + dot[3][lane] : 0.0;
        if (r < i) VtVt[i][r] = wsum;
        const double tv = tau * v;
#pragma unroll
        for (int m = 0; m < QRB; ++m)
            if (m > i) x[m] -= tv * readlane_f64(wsum, m);
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int m = 0; m < QRB; ++m) {
            if (m < nb) {
                A[(long long)(p0 + r) * W + p0 + m] = (r <= m) ? x[m] : 0.0;
                V[r + (long long)m * N] = (r == m) ? 1.0 : (r > m ? x[m] : 0.0);
            }
        }
        if (Vrows) {
            double* vr = Vrows + (long long)b * vr_stride + (long long)r * QRB;
#pragma unroll
            for (int m = 0; m < QRB; ++m)
                if (m < nb) vr[m] = (r == m) ? 1.0 : (r > m ? x[m] : 0.0);
        }
    }
    if (r < QRB) {
        // row j of T: T(j, k) = 0 for k < j, so the sums run over all k < i without a guard
        // (tau_i = 0 for i >= nb leaves those columns zero)
        const int j = r;
        double tr[QRB];
#pragma unroll
        for (int i = 0; i < QRB; ++i) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < i; k += 2) {
                const double2 vv = *reinterpret_cast<const double2*>(&VtVt[i][k]);
                s += tr[k] * vv.x;
                if (k + 1 < i) s += tr[k + 1] * vv.y;
            }
            const double t = (i == j) ? tau_s[i] : -tau_s[i] * s;
            tr[i] = (i < j) ? 0.0 : t;
        }
#pragma unroll
        for (int k = 0; k < QRB; ++k) Tg[j + (long long)k * QRB] = tr[k];
    }
}

// Fused trailing update of one panel: every trailing column a of the panel rows (right-hand
// sides included) becomes a - V T^T V^T a, i.e. a - V (a^T V T)^T.  One workgroup takes 32
// trailing columns of one chain and holds them in registers.  256 threads; wave w owns the
// 16-row tiles rt = w, w + 4, w + 8, w + 12 of the 32-column tile At (R x 32), loaded straight
// into f64 accumulator layout: item r of tile (rt, ct) holds At(16 rt + 4 r + l/16, 16 ct + l%16).
// That is also the A-fragment of At^T for the k-step rows 16 rt + 4 r .. + 3, so
//   1. each wave forms its partial W = At^T V (32 x 32) over its own rows, V read row-major;
//   2. the four partials are summed in LDS in a fixed order (the only LDS traffic);
//   3. wave w forms tile (w / 2, w % 2) of W2^T = T^T W^T into LDS;
//   4. At -= V W2^T on its own tiles (V read column-major: 16 consecutive rows per fragment, the
//      B-fragments of W2^T from LDS), and stores them.
// At is read and written exactly once; V and T come from L2.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void qr_update_kernel(double* __restrict__ bufs, long long stride, int N,
                                                            int W, int batch, int p0, int nb, int cb0, int ncb,
                                                            const double* __restrict__ Vs,
                                                            const double* __restrict__ Vrows, long long vr_stride,
                                                            const double* __restrict__ Ts)
{
    __shared__ double Wp[4][QRB][QRB + 1];
    typedef Mfma<double>::acc_t acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane >> 4, lc = lane & 15;
    const int R = N - p0, cols = W - (p0 + nb), nrt = (R + 15) >> 4;
    int b, cb;
    chain_tile(ncb, batch, b, cb);
    cb += cb0;
    const int c0 = cb * QRB;
    double* A = bufs + (long long)b * stride;
    const double* V = Vs + (long long)b * N * QRB;
    const double* Vr = Vrows + (long long)b * vr_stride;
    const double* T = Ts + (long long)b * QRB * QRB;
    const long long base = (long long)p0 * W + p0 + nb + c0;   // At(0, 0)

    acc_t at[4][2];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * (wave + 4 * q) + 4 * r + lr, col = 16 * ct + lc;
                at[q][ct][r] = (row < R && c0 + col < cols) ? A[base + (long long)row * W + col] : 0.0;
            }

    // 1. partial W(i, m) = sum over this wave's rows k of At(k, i) V(k, m)
    acc_t wp[2][2];
#pragma unroll
    for (int it = 0; it < 2; ++it)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) wp[it][mt] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (wave + 4 * q >= nrt) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 16 * (wave + 4 * q) + 4 * r + lr;
            double bv[2];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt) {
                const int m = 16 * mt + lc;
                bv[mt] = (k < R && m < nb) ? Vr[(long long)k * QRB + m] : 0.0;
            }
#pragma unroll
            for (int it = 0; it < 2; ++it)
#pragma unroll
                for (int mt = 0; mt < 2; ++mt) wp[it][mt] = Mfma<double>::mma(at[q][it][r], bv[mt], wp[it][mt]);
        }
    }
    // 2. W = sum of the four partials, in wave order
#pragma unroll
    for (int it = 0; it < 2; ++it)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) Wp[wave][16 * it + 4 * r + lr][16 * mt + lc] = wp[it][mt][r];
    __syncthreads();
    for (int e = tid; e < QRB * QRB; e += 256) {
        const int i = e >> 5, m = e & 31;
        Wp[0][i][m] = ((Wp[0][i][m] + Wp[1][i][m]) + Wp[2][i][m]) + Wp[3][i][m];
    }
    __syncthreads();
    // 3. W2^T(m, i) = sum_j T(j, m) W(i, j), tile (mt, it) = (wave / 2, wave % 2) by wave; stored
    //    negated in Wp[1] as W2n[m][i]
    {
        const int mt = wave >> 1, it = wave & 1;
        acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int j0 = 0; j0 < QRB; j0 += 4) {
            const int j = j0 + lr, m = 16 * mt + lc;
            const double av = (j < nb && m < nb) ? T[j + (long long)m * QRB] : 0.0;
            const double bv = Wp[0][16 * it + lc][j];
            acc = Mfma<double>::mma(av, bv, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) Wp[1][16 * mt + 4 * r + lr][16 * it + lc] = -acc[r];
    }
    __syncthreads();
    // 4. At -= V W2^T on this wave's tiles
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rt = wave + 4 * q;
        if (rt >= nrt) break;
        double av[2][4];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * rt + lc, m = 16 * mt + 4 * r + lr;
                av[mt][r] = (row < R && m < nb) ? V[row + (long long)m * N] : 0.0;
            }
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            acc_t acc = at[q][ct];
#pragma unroll
            for (int mt = 0; mt < 2; ++mt)
#pragma unroll
                for (int r = 0; r < 4; ++r)   // B(k = m, n = i) = -W2^T(16 mt + 4 r + l/16, 16 ct + l%16)
                    acc = Mfma<double>::mma(av[mt][r], Wp[1][16 * mt + 4 * r + lr][16 * ct + lc], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * rt + 4 * r + lr, col = 16 * ct + lc;
                if (row < R && c0 + col < cols) A[base + (long long)row * W + col] = acc[r];
            }
        }
    }
}

hipError_t launch_qr_blocked(double* buf, long long stride, int N, int W, int batch, double* Vs, double* Ts,
                             double* Wt, long long wt_stride, int tile, int update, int panel, hipStream_t s,
                             const LookAhead* la)
{
    // per chain: Vs N*QRB, Ts QRB*QRB, Wt 2 * wt_stride (Wt and Wt T; in QR_UPDATE_REG the first
    // holds the row-major copy of V, the second that of the look-ahead's parity-1 V), wt_stride >= W*QRB
    if (N > 256 || N < 1 || W <= N) return hipErrorInvalidValue;
    if (wt_stride < (long long)W * QRB) return hipErrorInvalidValue;
    double* Wt2 = Wt + (long long)batch * wt_stride;
    auto launch_panel = [&](int p0, int par, hipStream_t st) {
        const int nb = (N - p0 < QRB) ? N - p0 : QRB;
        double *V = par ? la->V2 : Vs, *T = par ? la->T2 : Ts, *Vr = par ? Wt2 : Wt;
        if (panel == QR_PANEL_1024)
            hipLaunchKernelGGL(qr_panel4_kernel, dim3(batch), dim3(1024), 0, st, buf, stride, N, W, p0, nb, V, T, Vr,
                               wt_stride);
        else
            hipLaunchKernelGGL(qr_panel_kernel, dim3(batch), dim3(256), 0, st, buf, stride, N, W, p0, nb, V, T, Vr,
                               wt_stride);
        return hipGetLastError();
    };
    auto launch_update = [&](int p0, int par, int cb0, int ncb, hipStream_t st) {
        const int nb = (N - p0 < QRB) ? N - p0 : QRB;
        double *V = par ? la->V2 : Vs, *T = par ? la->T2 : Ts, *Vr = par ? Wt2 : Wt;
        hipLaunchKernelGGL(qr_update_kernel, dim3(ncb * batch), dim3(256), 0, st, buf, stride, N, W, batch, p0, nb,
                           cb0, ncb, V, Vr, wt_stride, T);
        return hipGetLastError();
    };
    if (la && la->s2 && update == QR_UPDATE_REG) {
        // Look-ahead: the update of panel k first covers panel k + 1's columns (block 0); then panel
        // k + 1 (main stream) runs beside the rest of the update (second stream).  Panel buffers
        // alternate by parity, so panel k + 1 never overwrites what the rest of update k reads.
        const int np = (N + QRB - 1) / QRB;
        hipError_t e = launch_panel(0, 0, s);
        if (e != hipSuccess) return e;
        for (int k = 0; k < np; ++k) {
            const int p0 = k * QRB, nb = (N - p0 < QRB) ? N - p0 : QRB;
            const int cols = W - (p0 + nb), ncb = (cols + QRB - 1) / QRB;
            if (ncb <= 0) continue;
            if (k + 1 == np) {
                if ((e = launch_update(p0, k & 1, 0, ncb, s)) != hipSuccess) return e;
                continue;
            }
            if ((e = launch_update(p0, k & 1, 0, 1, s)) != hipSuccess) return e;
            if (ncb > 1) {
                if ((e = hipEventRecord(la->fork, s)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(la->s2, la->fork, 0)) != hipSuccess) return e;
            }
            // the panel is submitted first so that its workgroups take their slots before the update's
            if ((e = launch_panel(p0 + QRB, (k + 1) & 1, s)) != hipSuccess) return e;
            if (ncb > 1 && (e = launch_update(p0, k & 1, 1, ncb - 1, la->s2)) != hipSuccess) return e;
            if (ncb > 1) {
                if ((e = hipEventRecord(la->join, la->s2)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(s, la->join, 0)) != hipSuccess) return e;
            }
        }
        return hipSuccess;
    }
    for (int p0 = 0; p0 < N; p0 += QRB) {
        const int nb = (N - p0 < QRB) ? N - p0 : QRB;
        const int R = N - p0, cols = W - (p0 + nb);
        const bool reg = update == QR_UPDATE_REG;
        if (panel == QR_PANEL_1024)
            hipLaunchKernelGGL(qr_panel4_kernel, dim3(batch), dim3(1024), 0, s, buf, stride, N, W, p0, nb, Vs, Ts,
                               reg ? Wt : nullptr, wt_stride);
        else
            hipLaunchKernelGGL(qr_panel_kernel, dim3(batch), dim3(256), 0, s, buf, stride, N, W, p0, nb, Vs, Ts,
                               reg ? Wt : nullptr, wt_stride);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (cols <= 0) continue;
        if (reg) {
            hipLaunchKernelGGL(qr_update_kernel, dim3(((cols + QRB - 1) / QRB) * batch), dim3(256), 0, s, buf,
                               stride, N, W, batch, p0, nb, 0, (cols + QRB - 1) / QRB, Vs, Wt, wt_stride, Ts);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
            continue;
        }
        double* At = buf + (p0 + nb) + (long long)p0 * W;   // At^T: cols x R, column-major, ld W
        // Wt = At^T V
        GemmArgs<double> g1;
        g1.M = cols; g1.Nc = nb; g1.K = R;
        g1.A = At; g1.sA = stride; g1.lda = W; g1.tA = 0;
        g1.B = Vs; g1.sB = (long long)N * QRB; g1.ldb = N; g1.tB = 0;
        g1.C = Wt; g1.sC = wt_stride; g1.ldc = W; g1.tC = 0;
        g1.alpha = 1.0; g1.beta = 0.0; g1.batch = batch;
        e = launch_gemm<double>(g1, TILE_64x32, s);
        if (e != hipSuccess) return e;
        // Wt2 = Wt T
        GemmArgs<double> g2;
        g2.M = cols; g2.Nc = nb; g2.K = nb;
        g2.A = Wt; g2.sA = wt_stride; g2.lda = W; g2.tA = 0;
        g2.B = Ts; g2.sB = (long long)QRB * QRB; g2.ldb = QRB; g2.tB = 0;
        g2.C = Wt2; g2.sC = wt_stride; g2.ldc = W; g2.tC = 0;
        g2.alpha = 1.0; g2.beta = 0.0; g2.batch = batch;
        e = launch_gemm<double>(g2, TILE_64x32, s);
        if (e != hipSuccess) return e;
        // At^T -= Wt2 V^T
        GemmArgs<double> g3;
        g3.M = cols; g3.Nc = R; g3.K = nb;
        g3.A = Wt2; g3.sA = wt_stride; g3.lda = W; g3.tA = 0;
        g3.B = Vs; g3.sB = (long long)N * QRB; g3.ldb = N; g3.tB = 1;
        g3.C = At; g3.sC = stride; g3.ldc = W; g3.tC = 0;
        g3.alpha = -1.0; g3.beta = 1.0; g3.batch = batch;
        e = launch_gemm<double>(g3, tile, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------------------------
//  Blocked LU with partial pivoting of the leading N x N block of the row-major N x W system,
//  applied to the W - N trailing right-hand-side columns (the reference inverts this matrix with
//  Eigen's PartialPivLU, include/utils/stable_numerics.hpp:235, 304, 322).  Per panel of QRB columns:
//    lu_panel_kernel   4 waves per chain, thread r holds panel row r: per column an arg-max
//                      reduction (ties to the lower row, as LAPACK's idamax), the swap with the
//                      pivot row and the rank-1 elimination; writes L and U back, L column-major
//                      (Ls), the 32 x 32 inverse of the unit lower L11 (Li) and the row order of
//                      the panel rows (Ps).
//    lu_update_kernel  one workgroup per 32 trailing columns: loads the tile in the panel's row
//                      order, U12 = L11^{-1} A12 and A22 -= L21 U12 on MFMA.
//  On exit the leading block holds U (upper, with its diagonal) and the trailing columns hold
//  L^{-1} P rhs, so back-substitution with U finishes the solve.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int xchg_i32(int v, int lane, bool row16)
{
    const auto r = row16 ? __builtin_amdgcn_permlane16_swap((unsigned)v, (unsigned)v, false, false)
                         : __builtin_amdgcn_permlane32_swap((unsigned)v, (unsigned)v, false, false);
    const bool upper = row16 ? (lane & 16) : (lane & 32);
    return (int)(upper ? r[0] : r[1]);
}
__device__ __forceinline__ double xchg_f64(double v, int lane, bool row16)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)xchg_i32((int)(unsigned)u, lane, row16);
    const unsigned hi = (unsigned)xchg_i32((int)(unsigned)(u >> 32), lane, row16);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// larger value wins, equal values go to the smaller index
__device__ __forceinline__ void argmax_take(double& v, int& i, double ov, int oi)
{
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
}
// (value, index) arg-max over the 64 lanes; every lane ends with the result.  The pairings (quad
// perms, half-row and row mirrors, row-pair and half swaps) each join two groups that already
// agree, so after the six levels the whole wave agrees.
__device__ __forceinline__ void wave_argmax(double& v, int& i, int lane)
{
    argmax_take(v, i, dpp_f64<DPP_XOR1>(v), __builtin_amdgcn_mov_dpp(i, DPP_XOR1, 0xF, 0xF, false));
    argmax_take(v, i, dpp_f64<DPP_XOR2>(v), __builtin_amdgcn_mov_dpp(i, DPP_XOR2, 0xF, 0xF, false));
    argmax_take(v, i, dpp_f64<DPP_HALF_MIRROR>(v), __builtin_amdgcn_mov_dpp(i, DPP_HALF_MIRROR, 0xF, 0xF, false));
    argmax_take(v, i, dpp_f64<DPP_MIRROR>(v), __builtin_amdgcn_mov_dpp(i, DPP_MIRROR, 0xF, 0xF, false));
    argmax_take(v, i, xchg_f64(v, lane, true), xchg_i32(i, lane, true));
    argmax_take(v, i, xchg_f64(v, lane, false), xchg_i32(i, lane, false));
}

// One barrier per column: before it every wave publishes its arg-max (value, row) and that row
// (values and original index), and row i publishes itself; after it every thread picks the
// winning wave, row i takes the pivot row, the pivot's old position takes row i, and the rows
// below eliminate with the pivot row (read by lane j from LDS, then v_readlane to every lane).
// The published slots alternate between two buffers by column parity: a slot written for column
// i + 2 is past the barrier of column i + 1, which every reader of column i reached after its reads.
__global__ __launch_bounds__(256) void lu_panel_kernel(double* __restrict__ bufs, long long stride, int N, int W,
                                                       int p0, int nb, double* __restrict__ Ls,
                                                       double* __restrict__ Li, int* __restrict__ Ps,
                                                       uint32_t* __restrict__ health)
{
    __shared__ double redv[2][4];
    __shared__ int redi[2][4];
    __shared__ __align__(16) double cand[2][4][QRB];
    __shared__ int cando[2][4];
    __shared__ __align__(16) double irow[2][QRB];
    __shared__ int irowo[2];
    __shared__ double L11[QRB][QRB + 1];
    const int r = threadIdx.x, lane = r & 63, wave = r >> 6;
    const int b = blockIdx.x;
    const int R = N - p0;
    const bool live = r < R;
    double* A = bufs + (long long)b * stride;

    double x[QRB];
#pragma unroll
    for (int m = 0; m < QRB; ++m) x[m] = (live && m < nb) ? A[(long long)(p0 + r) * W + p0 + m] : 0.0;
    int org = r;
    bool zero_pivot = false;

#pragma clang loop unroll(full)
    for (int i = 0; i < QRB; ++i) {
        if (i >= nb) continue;   // uniform
        const int par = i & 1;
        double cv = (live && r >= i) ? fabs(x[i]) : -1.0;
        int ci = r;
        wave_argmax(cv, ci, lane);
        if (r == ci) {   // this lane holds its wave's candidate row
#pragma unroll
            for (int m = 0; m < QRB; m += 2) {
                double2 v; v.x = x[m]; v.y = x[m + 1];
                *reinterpret_cast<double2*>(&cand[par][wave][m]) = v;
            }
            cando[par][wave] = org;
        }
        if (lane == 0) { redv[par][wave] = cv; redi[par][wave] = ci; }
        if (r == i) {
#pragma unroll
            for (int m = 0; m < QRB; m += 2) {
                double2 v; v.x = x[m]; v.y = x[m + 1];
                *reinterpret_cast<double2*>(&irow[par][m]) = v;
            }
            irowo[par] = org;
        }
        __syncthreads();
        double bv = redv[par][0];
        int bi = redi[par][0], bw = 0;
#pragma unroll
        for (int w = 1; w < 4; ++w) {
            const double v = redv[par][w];
            const int ix = redi[par][w];
            if (v > bv || (v == bv && ix < bi)) { bv = v; bi = ix; bw = w; }
        }
        // pivot row to every lane: lane j reads entry j, v_readlane broadcasts (uniform values)
        const double pj = (lane < QRB) ? cand[par][bw][lane] : 0.0;
        if (r == i && bi != i) {
#pragma unroll
            for (int m = 0; m < QRB; ++m) x[m] = cand[par][bw][m];
            org = cando[par][bw];
        } else if (r == bi && bi != i) {
#pragma unroll
            for (int m = 0; m < QRB; ++m) x[m] = irow[par][m];
            org = irowo[par];
        }
        const double piv = readlane_f64(pj, i);
        zero_pivot |= !(piv != 0.0);
        const double l = (live && r > i) ? x[i] / piv : 0.0;
        if (r > i) x[i] = l;
#pragma unroll
        for (int m = 0; m < QRB; ++m)
            if (m > i) x[m] -= l * readlane_f64(pj, m);
    }
    if (r == 0 && zero_pivot) atomicOr(health + b, HEALTH_ZERO_PIVOT);
    if (live) {
#pragma unroll
        for (int m = 0; m < QRB; ++m)
            if (m < nb) {
                A[(long long)(p0 + r) * W + p0 + m] = x[m];
                Ls[(long long)b * N * QRB + r + (long long)m * N] = (r > m) ? x[m] : (r == m ? 1.0 : 0.0);
            }
        Ps[(long long)b * N + r] = org;
    }
    if (r < QRB) {
#pragma unroll
        for (int m = 0; m < QRB; ++m) L11[r][m] = (r > m && r < nb && m < nb) ? x[m] : 0.0;
    }
    __syncthreads();
    if (r < QRB) {
        // column j of L11^{-1} by forward substitution: y = 0 above j, y_j = 1, and the terms with
        // q < j vanish, so the sums need no guard; columns j >= nb stay zero
        const int j = r;
        double y[QRB];
#pragma unroll
        for (int m = 0; m < QRB; ++m) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < m; ++q) s += L11[m][q] * y[q];
            y[m] = (m < j || j >= nb) ? 0.0 : (m == j ? 1.0 : -s);
        }
#pragma unroll
        for (int m = 0; m < QRB; ++m) Li[(long long)b * QRB * QRB + m + (long long)j * QRB] = y[m];
    }
}

// Trailing update of one LU panel on a 32-column tile, 256 threads: wave w holds the 16-row tiles
// rt = w, w + 4, w + 8, w + 12 of the tile, rows loaded in the panel's row order (Ps) straight into
// f64 accumulator layout (item r of tile (rt, ct) = At(16 rt + 4 r + l/16, 16 ct + l%16)).  Waves 0
// and 1 hold the QRB panel rows A12: through LDS they form U12 = L11^{-1} A12, which every wave then
// reads as B-fragments for A22 -= L21 U12 on its own row tiles.  The tile is read and written once.
__global__ __launch_bounds__(256) void lu_update_kernel(double* __restrict__ bufs, long long stride, int N, int W,
                                                        int batch, int p0, int nb, int cb0, int ncb,
                                                        const double* __restrict__ Ls, const double* __restrict__ Li,
                                                        const int* __restrict__ Ps)
{
    __shared__ double S[2][QRB][QRB + 1];   // [0]: A12, [1]: U12
    typedef Mfma<double>::acc_t acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane >> 4, lc = lane & 15;
    const int R = N - p0, cols = W - (p0 + nb), nrt = (R + 15) >> 4;
    int b, cb;
    chain_tile(ncb, batch, b, cb);
    cb += cb0;
    const int c0 = cb * QRB;
    double* A = bufs + (long long)b * stride;
    const double* L = Ls + (long long)b * N * QRB;
    const double* Lv = Li + (long long)b * QRB * QRB;
    const int* P = Ps + (long long)b * N;
    const long long base = (long long)p0 * W + p0 + nb + c0;   // At(0, 0)

    acc_t at[4][2];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 16 * (wave + 4 * q) + 4 * r + lr;
            const int src = row < R ? P[row] : 0;
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                const int col = 16 * ct + lc;
                at[q][ct][r] = (row < R && c0 + col < cols) ? A[base + (long long)src * W + col] : 0.0;
            }
        }
    if (wave < 2) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) S[0][16 * wave + 4 * r + lr][16 * ct + lc] = at[0][ct][r];
    }
    __syncthreads();
    if (wave < 2) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
            for (int k0 = 0; k0 < QRB; k0 += 4)
                acc = Mfma<double>::mma(Lv[(16 * wave + lc) + (long long)(k0 + lr) * QRB], S[0][k0 + lr][16 * ct + lc], acc);
            at[0][ct] = acc;
#pragma unroll
            for (int r = 0; r < 4; ++r) S[1][16 * wave + 4 * r + lr][16 * ct + lc] = acc[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int rt = wave + 4 * q;
        if (rt < 2 || rt >= nrt) continue;
        const int arow = 16 * rt + lc;
        double av[QRB / 4];
#pragma unroll
        for (int s = 0; s < QRB / 4; ++s) av[s] = (arow < R) ? -L[arow + (long long)(4 * s + lr) * N] : 0.0;
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            acc_t acc = at[q][ct];
#pragma unroll
            for (int s = 0; s < QRB / 4; ++s) acc = Mfma<double>::mma(av[s], S[1][4 * s + lr][16 * ct + lc], acc);
            at[q][ct] = acc;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (wave + 4 * q >= nrt) continue;
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * (wave + 4 * q) + 4 * r + lr, col = 16 * ct + lc;
                if (row < R && c0 + col < cols) A[base + (long long)row * W + col] = at[q][ct][r];
            }
    }
}

hipError_t launch_lu_blocked(double* buf, long long stride, int N, int W, int batch, double* Ls, double* Li, int* Ps,
                             uint32_t* health, hipStream_t s, const LookAhead* la)
{
    if (N > 256 || N < 1 || W <= N) return hipErrorInvalidValue;
    const bool ahead = la && la->s2;
    auto launch_panel = [&](int p0, int par, hipStream_t st) {
        const int nb = (N - p0 < QRB) ? N - p0 : QRB;
        hipLaunchKernelGGL(lu_panel_kernel, dim3(batch), dim3(256), 0, st, buf, stride, N, W, p0, nb,
                           par ? la->V2 : Ls, par ? la->T2 : Li, par ? la->P2 : Ps, health);
        return hipGetLastError();
    };
    auto launch_update = [&](int p0, int par, int cb0, int ncb, hipStream_t st) {
        const int nb = (N - p0 < QRB) ? N - p0 : QRB;
        hipLaunchKernelGGL(lu_update_kernel, dim3(ncb * batch), dim3(256), 0, st, buf, stride, N, W, batch, p0, nb,
                           cb0, ncb, par ? la->V2 : Ls, par ? la->T2 : Li, par ? la->P2 : Ps);
        return hipGetLastError();
    };
    const int np = (N + QRB - 1) / QRB;
    hipError_t e = launch_panel(0, 0, s);
    if (e != hipSuccess) return e;
    for (int k = 0; k < np; ++k) {
        const int p0 = k * QRB, nb = (N - p0 < QRB) ? N - p0 : QRB;
        const int cols = W - (p0 + nb), ncb = (cols + QRB - 1) / QRB;
        const int par = ahead ? (k & 1) : 0, npar = ahead ? ((k + 1) & 1) : 0;
        if (!ahead || k + 1 == np) {
            // in order: the whole update of panel k, then panel k + 1
            if (ncb > 0 && (e = launch_update(p0, par, 0, ncb, s)) != hipSuccess) return e;
            if (k + 1 < np && (e = launch_panel(p0 + QRB, npar, s)) != hipSuccess) return e;
            continue;
        }
        // look-ahead, as in launch_qr_blocked: panel k + 1's columns first, then panel k + 1 beside
        // the rest of the update (panel buffers and row orders alternate by parity)
        if ((e = launch_update(p0, par, 0, 1, s)) != hipSuccess) return e;
        if (ncb > 1) {
            if ((e = hipEventRecord(la->fork, s)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(la->s2, la->fork, 0)) != hipSuccess) return e;
        }
        if ((e = launch_panel(p0 + QRB, npar, s)) != hipSuccess) return e;
        if (ncb > 1 && (e = launch_update(p0, par, 1, ncb - 1, la->s2)) != hipSuccess) return e;
        if (ncb > 1) {
            if ((e = hipEventRecord(la->join, la->s2)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(s, la->join, 0)) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------------------------
//  Blocked back-substitution X = R^{-1} Z for the row-major system (R in columns [0, N), Z in
//  columns [c0, c0+P), solved in place), 32-row blocks from the bottom:
//    trsm_diag_kernel  X_k = R_kk^{-1} Z_k, one thread per right-hand side, R_kk in LDS;
//    GEMM (MFMA)       Z_{0:k} -= R_{0:k, k} X_k, as Z^T_{:, 0:k} -= X_k^T R_{0:k, k}^T.
// ---------------------------------------------------------------------------------------------
#define TRB 32

__global__ __launch_bounds__(256) void trsm_diag_kernel(double* __restrict__ bufs, long long stride, int W, int c0,
                                                        int P, int k0, int nbk)
{
    __shared__ double Rk[TRB][TRB + 1];
    double* A = bufs + (long long)blockIdx.y * stride;
    for (int e = threadIdx.x; e < TRB * TRB; e += 256) {
        const int m = e / TRB, q = e % TRB;
        Rk[m][q] = (m < nbk && q < nbk) ? A[(long long)(k0 + m) * W + k0 + q] : 0.0;
    }
    __syncthreads();
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= P) return;
    double* Z = A + c0 + j;
    double x[TRB];
#pragma unroll
    for (int m = 0; m < TRB; ++m) x[m] = (m < nbk) ? Z[(long long)(k0 + m) * W] : 0.0;
#pragma unroll
    for (int m = TRB - 1; m >= 0; --m) {
        if (m < nbk) {
            double v = x[m];
#pragma unroll
            for (int q = m + 1; q < TRB; ++q)
                if (q < nbk) v -= Rk[m][q] * x[q];
            x[m] = v / Rk[m][m];
        }
    }
#pragma unroll
    for (int m = 0; m < TRB; ++m)
        if (m < nbk) Z[(long long)(k0 + m) * W] = x[m];
}

hipError_t launch_trsm_blocked(double* buf, long long stride, int N, int W, int c0, int P, int batch, int tile,
                               hipStream_t s)
{
    if (N < 1 || P < 1 || c0 < N || c0 + P > W) return hipErrorInvalidValue;
    const int nblk = (N + TRB - 1) / TRB;
    for (int kb = nblk - 1; kb >= 0; --kb) {
        const int k0 = kb * TRB, nbk = (N - k0 < TRB) ? N - k0 : TRB;
        hipLaunchKernelGGL(trsm_diag_kernel, dim3((P + 255) / 256, batch), dim3(256), 0, s, buf, stride, W, c0, P, k0,
                           nbk);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (k0 == 0) break;
        GemmArgs<double> g;
        g.M = P; g.Nc = k0; g.K = nbk;
        g.A = buf + c0 + (long long)k0 * W; g.sA = stride; g.lda = W; g.tA = 0;
        g.B = buf + k0; g.sB = stride; g.ldb = W; g.tB = 0;
        g.C = buf + c0; g.sC = stride; g.ldc = W; g.tC = 0;
        g.alpha = -1.0; g.beta = 1.0; g.batch = batch;
        e = launch_gemm<double>(g, tile, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------------------------
//  Upper triangular solve (DQMC_TRSM=v1), thread-per-right-hand-side column, 16-row blocks from
//  the bottom.
// ---------------------------------------------------------------------------------------------
#define TRSM_BR 16
__global__ __launch_bounds__(256) void trsm_v1_kernel(double* __restrict__ bufs, long long stride, int N, int W,
                                                   int c0, int P)
{
    double* A = bufs + (long long)blockIdx.y * stride;
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= P) return;
    double* Z = A + c0 + col;
    for (int iend = N; iend > 0; iend -= TRSM_BR) {
        const int i0 = iend - TRSM_BR > 0 ? iend - TRSM_BR : 0;
        const int bs = iend - i0;
        double x[TRSM_BR];
#pragma unroll
        for (int m = 0; m < TRSM_BR; ++m) x[m] = (m < bs) ? Z[(long long)(i0 + m) * W] : 0.0;
#pragma unroll
        for (int m = TRSM_BR - 1; m >= 0; --m) {
            if (m < bs) {
                double s = x[m];
#pragma unroll
                for (int q = m + 1; q < TRSM_BR; ++q)
                    if (q < bs) s -= A[(long long)(i0 + m) * W + i0 + q] * x[q];
                x[m] = s / A[(long long)(i0 + m) * W + i0 + m];
            }
        }
#pragma unroll
        for (int m = 0; m < TRSM_BR; ++m)
            if (m < bs) Z[(long long)(i0 + m) * W] = x[m];
        for (int k = 0; k < i0; ++k) {
            const double* Rk = A + (long long)k * W + i0;
            double s = Z[(long long)k * W];
#pragma unroll
            for (int m = 0; m < TRSM_BR; ++m)
                if (m < bs) s -= Rk[m] * x[m];
            Z[(long long)k * W] = s;
        }
    }
}

hipError_t launch_trsm_upper_v1(double* buf, long long stride, int N, int W, int c0, int P, int batch,
                             hipStream_t s)
{
    dim3 grid((P + 255) / 256, batch);
    hipLaunchKernelGGL(trsm_v1_kernel, grid, dim3(256), 0, s, buf, stride, N, W, c0, P);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
//  UDV helpers
// ---------------------------------------------------------------------------------------------
// Column norms of M (column-major, ld N) and the descending-norm order (ties to the lower index,
// a stable sort): one workgroup per chain, wave w reduces the column groups 8 w .. 8 w + 7,
// 8 w + 32 .., with its lanes reading consecutive rows.
__global__ __launch_bounds__(256) void udv_rank_kernel(const double* __restrict__ Ms, long long sM,
                                                       int* __restrict__ perms, int N)
{
    __shared__ double nrm[1024];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* M = Ms + (long long)b * sM;
    int* perm = perms + (long long)b * N;
    // 8 columns per step, so that 8 x ceil(N / 64) loads per lane are in flight together
    for (int j0 = wave * 8; j0 < N; j0 += 32) {
        double s[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) s[c] = 0.0;
        for (int r = lane; r < N; r += 64) {
            double v[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) v[c] = (j0 + c < N) ? M[r + (long long)(j0 + c) * N] : 0.0;
#pragma unroll
            for (int c = 0; c < 8; ++c) s[c] += v[c] * v[c];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const double t = wave_sum_dpp(s[c]);
            if (lane == 0 && j0 + c < N) nrm[j0 + c] = t;
        }
    }
    __syncthreads();
    for (int j = tid; j < N; j += 256) {
        const double nj = nrm[j];
        int rank = 0;
        for (int i = 0; i < N; ++i) {
            const double ni = nrm[i];
            rank += (ni > nj) || (ni == nj && i < j);
        }
        perm[rank] = j;
    }
}

// buf[r][j] = M(r, perm[j]) and buf[r][N + j] = delta_rj, through a 32 x 32 LDS tile so that the
// column reads of M and the row writes of buf are both contiguous; grid (tiles, chains).
__global__ __launch_bounds__(256) void udv_copy_kernel(const double* __restrict__ Ms, long long sM,
                                                       double* __restrict__ bufs, long long sbuf,
                                                       const int* __restrict__ perms, int N, int W)
{
    __shared__ double t[32][33];
    const int b = blockIdx.y, nt = (N + 31) / 32;
    const int r0 = (blockIdx.x % nt) * 32, j0 = (blockIdx.x / nt) * 32;
    const double* M = Ms + (long long)b * sM;
    double* buf = bufs + (long long)b * sbuf;
    const int* perm = perms + (long long)b * N;
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int rr = e & 31, jj = e >> 5;
        const int r = r0 + rr, j = j0 + jj;
        t[jj][rr] = (r < N && j < N) ? M[r + (long long)perm[j] * N] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int jj = e & 31, rr = e >> 5;
        const int r = r0 + rr, j = j0 + jj;
        if (r < N && j < N) {
            buf[(long long)r * W + j] = t[jj][rr];
            buf[(long long)r * W + N + j] = (r == j) ? 1.0 : 0.0;
        }
    }
}

hipError_t launch_udv_prepare(const double* M, long long sM, double* buf, long long sbuf, int W, int* perm, int N,
                              int batch, hipStream_t s)
{
    if (N > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(udv_rank_kernel, dim3(batch), dim3(256), 0, s, M, sM, perm, N);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nt = (N + 31) / 32;
    hipLaunchKernelGGL(udv_copy_kernel, dim3(nt * nt, batch), dim3(256), 0, s, M, sM, buf, sbuf, perm, N, W);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void udv_finish_kernel(const double* __restrict__ bufs, long long sbuf,
                                                         const int* __restrict__ perms, double* __restrict__ Us,
                                                         long long sU, double* __restrict__ Ds, long long sD,
                                                         double* __restrict__ Tts, long long sTt, int N, int W,
                                                         uint32_t* health)
{
    const int b = blockIdx.y;
    const double* buf = bufs + (long long)b * sbuf;
    const int* perm = perms + (long long)b * N;
    double* U = Us + (long long)b * sU;
    double* D = Ds + (long long)b * sD;
    double* Tt = Tts + (long long)b * sTt;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)N * N) return;
    const int i = (int)(e / N), j = (int)(e % N);
    // U column-major: U(j, i) = Q(j, i) = (Q^T)(i, j) = buf[i][N + j]
    U[j + (long long)i * N] = buf[(long long)i * W + N + j];
    const double d = fabs(buf[(long long)i * W + i]);
    if (j == 0) {
        D[i] = d;
        // a zero or non-finite |R_ii| is a breakdown of the factorization
        if (!(d > 0.0) || !isfinite(d)) atomicOr(health + b, HEALTH_ZERO_PIVOT);
    }
    // Tt(perm[j], i) = R(i, j) / D_i for j >= i
    Tt[perm[j] + (long long)i * N] = (j >= i) ? buf[(long long)i * W + j] / d : 0.0;
}

hipError_t launch_udv_finish(const double* buf, long long sbuf, int W, const int* perm, double* U, long long sU,
                             double* D, long long sD, double* Tt, long long sTt, int N, uint32_t* health, int batch,
                             hipStream_t s)
{
    dim3 grid((unsigned)(((long long)N * N + 255) / 256), batch);
    hipLaunchKernelGGL(udv_finish_kernel, grid, dim3(256), 0, s, buf, sbuf, perm, U, sU, D, sD, Tt, sTt, N, W, health);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void build_00bb_kernel(const double* __restrict__ us, long long su,
                                                         const double* __restrict__ vs, long long sv,
                                                         const double* __restrict__ ds, long long sd,
                                                         double* __restrict__ bufs, long long sbuf, int N, int W)
{
    const int b = blockIdx.y;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)N * N) return;
    const int i = (int)(e / N), j = (int)(e % N);
    const double d = ds[(long long)b * sd + i];
    const double sl = 1.0 / fmax(1.0, d), sr = fmin(1.0, d);
    const double uji = us[(long long)b * su + j + (long long)i * N];
    const double vji = vs[(long long)b * sv + j + (long long)i * N];
    double* buf = bufs + (long long)b * sbuf;
    buf[(long long)i * W + j] = sl * uji + sr * vji;
    buf[(long long)i * W + N + j] = sl * uji;
}

hipError_t launch_build_00bb(const double* u, long long su, const double* v, long long sv, const double* d,
                             long long sd, double* buf, long long sbuf, int N, int W, int batch, hipStream_t s)
{
    dim3 grid((unsigned)(((long long)N * N + 255) / 256), batch);
    hipLaunchKernelGGL(build_00bb_kernel, grid, dim3(256), 0, s, u, su, v, sv, d, sd, buf, sbuf, N, W);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void build_rhs_kernel(const double* __restrict__ Xs, long long sX, DiagArg d,
                                                        double alpha, double* __restrict__ bufs, long long sbuf,
                                                        int N, int W, int c0)
{
    const int b = blockIdx.y;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)N * N) return;
    const int r = (int)(e / N), j = (int)(e % N);
    const double sc = alpha * diag_value(d, b, r);
    bufs[(long long)b * sbuf + (long long)r * W + c0 + j] = sc * Xs[(long long)b * sX + j + (long long)r * N];
}

hipError_t launch_build_rhs(const double* X, long long sX, DiagArg d, double alpha, double* buf, long long sbuf,
                            int N, int W, int c0, int batch, hipStream_t s)
{
    dim3 grid((unsigned)(((long long)N * N + 255) / 256), batch);
    hipLaunchKernelGGL(build_rhs_kernel, grid, dim3(256), 0, s, X, sX, d, alpha, buf, sbuf, N, W, c0);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void extract_kernel(const double* __restrict__ bufs, long long sbuf, int W,
                                                      int c0, double* __restrict__ outs, long long so, int N,
                                                      int transpose)
{
    const int b = blockIdx.y;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)N * N) return;
    const int r = (int)(e / N), j = (int)(e % N);
    const double v = bufs[(long long)b * sbuf + (long long)r * W + c0 + j];
    double* out = outs + (long long)b * so;
    if (transpose) out[r + (long long)j * N] = v;
    else out[j + (long long)r * N] = v;
}

hipError_t launch_extract(const double* buf, long long sbuf, int W, int c0, double* out, long long so, int N,
                          int transpose, int batch, hipStream_t s)
{
    dim3 grid((unsigned)(((long long)N * N + 255) / 256), batch);
    hipLaunchKernelGGL(extract_kernel, grid, dim3(256), 0, s, buf, sbuf, W, c0, out, so, N, transpose);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
//  Error reductions and commit (one workgroup per chain, single writer per chain).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ double block_max(double v, double* red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, red[w]);
    __syncthreads();
    return r;
}

// Stage 1: workgroup p of chain b takes entries [p * span, (p + 1) * span) of the N x N matrices,
// commits them and writes its partial (signed max, max-abs, non-finite flag) to part[b][p].
// Stage 2: one thread per chain folds the parts in order into the running maxima and the health
// flags (no atomics: the same bits on every run).
#define CMP_PARTS 16
__global__ __launch_bounds__(256) void compare_commit_kernel(const double* __restrict__ Gnew, double* __restrict__ G64,
                                                             float* __restrict__ G32, long long sG, int N, int span,
                                                             double* __restrict__ part)
{
    __shared__ double red[4];
    const int p = blockIdx.x, b = blockIdx.y, np = gridDim.x;
    const double* gn = Gnew + (long long)b * sG;
    double* g64 = G64 + (long long)b * sG;
    float* g32 = G32 ? G32 + (long long)b * sG : nullptr;
    const long long e0 = (long long)p * span, e1 = min((long long)N * N, e0 + span);
    double es = -DBL_MAX, ea = 0.0;
    int bad = 0;
    for (long long e = e0 + threadIdx.x; e < e1; e += 256) {
        const double v = gn[e];
        const double cur = g32 ? (double)g32[e] : g64[e];
        const double d = v - cur;
        es = fmax(es, d);
        ea = fmax(ea, fabs(d));
        bad |= !isfinite(v) || !isfinite(cur);
        g64[e] = v;
        if (g32) g32[e] = (float)v;
    }
    es = block_max(es, red);
    ea = block_max(ea, red);
    const double nanflag = block_max(bad ? 1.0 : 0.0, red);
    if (threadIdx.x == 0) {
        double* q = part + ((long long)b * np + p) * 3;
        q[0] = es; q[1] = ea; q[2] = nanflag;
    }
}

__global__ void compare_fold_kernel(const double* __restrict__ part, int np, double* err_signed, double* err_abs,
                                    const double* __restrict__ bound, uint32_t* health, int batch)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    double es = -DBL_MAX, ea = 0.0, nanflag = 0.0;
    for (int p = 0; p < np; ++p) {
        const double* q = part + ((long long)b * np + p) * 3;
        es = fmax(es, q[0]);
        ea = fmax(ea, q[1]);
        nanflag = fmax(nanflag, q[2]);
    }
    if (es > err_signed[b]) err_signed[b] = es;
    if (ea > err_abs[b] || nanflag > 0.0) err_abs[b] = nanflag > 0.0 ? INFINITY : fmax(err_abs[b], ea);
    uint32_t h = health[b];
    if (nanflag > 0.0) h |= HEALTH_NONFINITE_G;
    const double lim = *bound;
    if (lim > 0.0 && ea > lim) h |= HEALTH_WATCHDOG;
    health[b] = h;
}

hipError_t launch_compare_commit(const double* Gnew, double* G64, float* G32, long long sG, int N,
                                 double* err_signed, double* err_abs, const double* bound, uint32_t* health,
                                 double* part, int batch, hipStream_t s)
{
    const long long nn = (long long)N * N;
    int np = (int)((nn + 4095) / 4096);
    if (np > CMP_PARTS) np = CMP_PARTS;
    const int span = (int)((nn + np - 1) / np);
    hipLaunchKernelGGL(compare_commit_kernel, dim3(np, batch), dim3(256), 0, s, Gnew, G64, G32, sG, N, span, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(compare_fold_kernel, dim3((batch + 63) / 64), dim3(64), 0, s, part, np, err_signed, err_abs,
                       bound, health, batch);
    return hipGetLastError();
}

// Flags chains whose N x N matrix holds a non-finite entry (one workgroup per chain).
__global__ __launch_bounds__(256) void finite_check_kernel(const double* __restrict__ A, long long sA, int N,
                                                           uint32_t* health)
{
    __shared__ double red[4];
    const int b = blockIdx.x;
    int bad = 0;
    for (long long e = threadIdx.x; e < (long long)N * N; e += 256) bad |= !isfinite(A[(long long)b * sA + e]);
    const double f = block_max(bad ? 1.0 : 0.0, red);
    if (threadIdx.x == 0 && f > 0.0) health[b] |= HEALTH_NONFINITE_G;
}

hipError_t launch_finite_check(const double* A, long long sA, int N, uint32_t* health, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(finite_check_kernel, dim3(batch), dim3(256), 0, s, A, sA, N, health);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void compare_only_kernel(const double* __restrict__ Gnew,
                                                           const double* __restrict__ Gcur, long long sG, int N,
                                                           double* err_signed, uint32_t* health)
{
    __shared__ double red[4];
    const int b = blockIdx.x;
    double es = -DBL_MAX;
    int bad = 0;
    for (long long e = threadIdx.x; e < (long long)N * N; e += 256) {
        const double v = Gnew[(long long)b * sG + e], cur = Gcur[(long long)b * sG + e];
        es = fmax(es, v - cur);
        bad |= !isfinite(v) || !isfinite(cur);
    }
    es = block_max(es, red);
    const double f = block_max(bad ? 1.0 : 0.0, red);
    if (threadIdx.x == 0) {
        if (es > err_signed[b]) err_signed[b] = es;
        if (f > 0.0) health[b] |= HEALTH_NONFINITE_G;
    }
}

hipError_t launch_compare_only(const double* Gnew, const double* Gcur, long long sG, int N, double* err_signed,
                               uint32_t* health, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(compare_only_kernel, dim3(batch), dim3(256), 0, s, Gnew, Gcur, sG, N, err_signed, health);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
//  Single-site Sherman-Morrison update (Model::HubbardAttractiveU::update_green_function,
//  src/hubbard.cpp:132-139): G(r, c) -= (gamma G(r, i)) (delta_ic - G(i, c)) with
//  gamma = delta / (1 + (1 - G_ii) delta) and delta = e^{-2 alpha s(t,i)} - 1, all from the G before
//  the update.  One workgroup per chain stages column i and row i in LDS first.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void site_update_kernel(double* __restrict__ G, int N, int i, double delta)
{
#pragma clang fp contract(off)
    __shared__ double col[1024], row[1024];
    for (int r = threadIdx.x; r < N; r += 256) {
        col[r] = G[r + (long long)i * N];
        row[r] = G[i + (long long)r * N];
    }
    __syncthreads();
    const double gamma = delta / (1.0 + (1.0 - col[i]) * delta);
    for (long long e = threadIdx.x; e < (long long)N * N; e += 256) {
        const int r = (int)(e % N), c = (int)(e / N);
        const double t = (gamma * col[r]) * ((c == i ? 1.0 : 0.0) - row[c]);
        G[e] = G[e] - t;
    }
}
hipError_t launch_site_update(double* G, int N, int i, double delta, hipStream_t s)
{
    if (N > 1024 || i < 0 || i >= N) return hipErrorInvalidValue;
    hipLaunchKernelGGL(site_update_kernel, dim3(1), dim3(256), 0, s, G, N, i, delta);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
//  Elementwise helpers
// ---------------------------------------------------------------------------------------------
__global__ void set_identity_kernel(double* A, long long s, int N, double dv)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)N * N) return;
    const int i = (int)(e % N), j = (int)(e / N);
    A[(long long)blockIdx.y * s + e] = (i == j) ? dv : 0.0;
}
hipError_t launch_set_identity(double* A, long long s, int N, double dv, int batch, hipStream_t st)
{
    dim3 grid((unsigned)(((long long)N * N + 255) / 256), batch);
    hipLaunchKernelGGL(set_identity_kernel, grid, dim3(256), 0, st, A, s, N, dv);
    return hipGetLastError();
}

template <typename S, typename D>
__global__ void copy_kernel(const S* src, long long ss, D* dst, long long sd, long long n)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    dst[(long long)blockIdx.y * sd + e] = (D)src[(long long)blockIdx.y * ss + e];
}
hipError_t launch_copy_d(const double* src, long long ss, double* dst, long long sd, long long n, int batch,
                         hipStream_t st)
{
    dim3 grid((unsigned)((n + 255) / 256), batch);
    hipLaunchKernelGGL((copy_kernel<double, double>), grid, dim3(256), 0, st, src, ss, dst, sd, n);
    return hipGetLastError();
}
hipError_t launch_copy_f2d(const float* src, long long ss, double* dst, long long sd, long long n, int batch,
                           hipStream_t st)
{
    dim3 grid((unsigned)((n + 255) / 256), batch);
    hipLaunchKernelGGL((copy_kernel<float, double>), grid, dim3(256), 0, st, src, ss, dst, sd, n);
    return hipGetLastError();
}
hipError_t launch_copy_d2f(const double* src, long long ss, float* dst, long long sd, long long n, int batch,
                           hipStream_t st)
{
    dim3 grid((unsigned)((n + 255) / 256), batch);
    hipLaunchKernelGGL((copy_kernel<double, float>), grid, dim3(256), 0, st, src, ss, dst, sd, n);
    return hipGetLastError();
}

__global__ void axpi_kernel(const double* src, long long ss, double* dst, long long sd, double a, double bb, int N)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)N * N) return;
    const int i = (int)(e % N), j = (int)(e / N);
    dst[(long long)blockIdx.y * sd + e] = a * src[(long long)blockIdx.y * ss + e] + (i == j ? bb : 0.0);
}
hipError_t launch_axpi(const double* src, long long ss, double* dst, long long sd, double a, double b, int N,
                       int batch, hipStream_t st)
{
    dim3 grid((unsigned)(((long long)N * N + 255) / 256), batch);
    hipLaunchKernelGGL(axpi_kernel, grid, dim3(256), 0, st, src, ss, dst, sd, a, b, N);
    return hipGetLastError();
}

__global__ void fill_kernel(double* dst, long long n, double v)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) dst[e] = v;
}
hipError_t launch_fill_d(double* dst, long long n, double v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, n, v);
    return hipGetLastError();
}

__global__ void fill_u32_kernel(uint32_t* dst, long long n, uint32_t v)
{
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e < n) dst[e] = v;
}
hipError_t launch_fill_u32(uint32_t* dst, long long n, uint32_t v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, n, v);
    return hipGetLastError();
}

}  // namespace dqmc

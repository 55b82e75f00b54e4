// dqmc_kernels.hip -- gfx950 kernels of the DQMC sweep engine.
//
//   gemm_kernel ........ batched GEMM on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32)
//                        with fused diagonal scalings: the slice wraps G <- B G B^{-1}, the
//                        B-chain products and the UDV/recompose products.
//   metropolis_kernel .. one workgroup per chain walks the N sites of a slice: FP64 ratio
//                        from G_ii, minstd_rand0 uniforms by jump-ahead, delayed rank-k
//                        Sherman-Morrison updates flushed into G inside the workgroup.
//   qr_kernel .......... blocked Householder QR (panels in LDS) of a row-major N x N block,
//                        applying Q^T to trailing right-hand-side columns.
//   trsm_kernel ........ X = R^{-1} Z, row blocks of 16 held in registers.
//   small helpers ...... UDV pre-pivot / finish, system builders, error reductions, copies.
#include "dqmc_kernels.h"

#include <float.h>

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

    const int b = blockIdx.z;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
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

    for (int k0 = 0; k0 < g.K; k0 += BK) {
#pragma unroll
        for (int e0 = 0; e0 < BM * BK; e0 += 256) {
            const int e = e0 + tid;
            int i, kk;
            if (!g.tA) { i = e % BM; kk = e / BM; } else { kk = e % BK; i = e / BK; }
            const int gi = m0 + i, gk = k0 + kk;
            T v = T(0);
            if (gi < g.M && gk < g.K) {
                v = g.tA ? A[gk + (long long)gi * g.lda] : A[gi + (long long)gk * g.lda];
                if (g.innd.mode != DIAG_NONE) v = T((double)v * diag_value(g.innd, b, gk));
            }
            As[kk][i] = v;
        }
#pragma unroll
        for (int e0 = 0; e0 < BN * BK; e0 += 256) {
            const int e = e0 + tid;
            int j, kk;
            if (!g.tB) { kk = e % BK; j = e / BK; } else { j = e % BN; kk = e / BN; }
            const int gj = n0 + j, gk = k0 + kk;
            T v = T(0);
            if (gj < g.Nc && gk < g.K) v = g.tB ? B[gj + (long long)gk * g.ldb] : B[gk + (long long)gj * g.ldb];
            Bs[kk][j] = v;
        }
        __syncthreads();
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

template <typename T, int BM, int BN>
static hipError_t launch_gemm_tile(const GemmArgs<T>& a, hipStream_t s)
{
    dim3 grid((a.M + BM - 1) / BM, (a.Nc + BN - 1) / BN, a.batch);
    hipLaunchKernelGGL((gemm_kernel<T, BM, BN, 16>), grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_gemm(const GemmArgs<T>& a, int tile, hipStream_t s)
{
    if (a.batch <= 0 || a.M <= 0 || a.Nc <= 0) return hipSuccess;
    switch (tile) {
        case TILE_32x32: return launch_gemm_tile<T, 32, 32>(a, s);
        case TILE_64x32: return launch_gemm_tile<T, 64, 32>(a, s);
        default: return launch_gemm_tile<T, 64, 64>(a, s);
    }
}
template hipError_t launch_gemm<double>(const GemmArgs<double>&, int, hipStream_t);
template hipError_t launch_gemm<float>(const GemmArgs<float>&, int, hipStream_t);

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

template <typename T, int KK>
__device__ __forceinline__ void flush_rows(T* __restrict__ G, int N, const T* __restrict__ X,
                                           const T* __restrict__ Y, double* diag0, int tid, int nth)
{
    for (int j = tid; j < N; j += nth) {
        T xr[KK];
#pragma unroll
        for (int m = 0; m < KK; ++m) xr[m] = X[m * N + j];
        for (int c = 0; c < N; ++c) {
            T acc = T(0);
#pragma unroll
            for (int m = 0; m < KK; ++m) acc += xr[m] * Y[m * N + c];
            const T gn = G[j + (long long)c * N] + acc;
            G[j + (long long)c * N] = gn;
            if (c == j) diag0[j] = (double)gn;
        }
    }
}

template <typename T>
__device__ void flush_pending(T* G, int N, T* X, T* Y, double* diag0, double* dcorr, int kc, int kmax, int tid,
                              int nth)
{
    // pad the pending terms with zeros up to the unrolled width
    const int kk = kmax == 1 ? 1 : ((kc + 7) / 8) * 8;
    for (int m = kc; m < kk; ++m)
        for (int j = tid; j < N; j += nth) { X[m * N + j] = T(0); Y[m * N + j] = T(0); }
    __syncthreads();
    switch (kk) {
        case 1: flush_rows<T, 1>(G, N, X, Y, diag0, tid, nth); break;
        case 8: flush_rows<T, 8>(G, N, X, Y, diag0, tid, nth); break;
        case 16: flush_rows<T, 16>(G, N, X, Y, diag0, tid, nth); break;
        case 24: flush_rows<T, 24>(G, N, X, Y, diag0, tid, nth); break;
        case 32: flush_rows<T, 32>(G, N, X, Y, diag0, tid, nth); break;
        case 40: flush_rows<T, 40>(G, N, X, Y, diag0, tid, nth); break;
        case 48: flush_rows<T, 48>(G, N, X, Y, diag0, tid, nth); break;
        case 56: flush_rows<T, 56>(G, N, X, Y, diag0, tid, nth); break;
        default: flush_rows<T, 64>(G, N, X, Y, diag0, tid, nth); break;
    }
    for (int j = tid; j < N; j += nth) dcorr[j] = 0.0;
    __syncthreads();
}

__device__ __forceinline__ double accept_prob(double gii, double delta, double* xout)
{
#pragma clang fp contract(off)
    const double x = 1.0 + (1.0 - gii) * delta;
    const double r = x * x / (1.0 + delta);
    *xout = x;
    return fmin(1.0, fabs(r));
}

template <typename T>
__global__ __launch_bounds__(256) void metropolis_kernel(MetroArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int N = a.N, kmax = a.kmax;
    const int b = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
    const int kal = kmax == 1 ? 1 : ((kmax + 7) / 8) * 8;   // rows of X / Y (flushes pad to 8)
    T* X = reinterpret_cast<T*>(smem);
    T* Y = X + (size_t)kal * N;
    double* dcorr = reinterpret_cast<double*>(Y + (size_t)kal * N);
    double* diag0 = dcorr + N;
    double* unif = diag0 + N;
    int* sl = reinterpret_cast<int*>(unif + N);
    uint32_t* mask = reinterpret_cast<uint32_t*>(sl + N);

    T* G = a.G + (long long)b * a.sG;
    int8_t* f = a.fields + ((long long)b * a.L + a.t) * N;
    const uint64_t x0 = a.rng[b];
    const int nw = (N + 31) / 32;

    for (int j = tid; j < N; j += nth) {
        sl[j] = f[j];
        diag0[j] = (double)G[j + (long long)j * N];
        dcorr[j] = 0.0;
        const uint64_t x1 = (x0 * a.powtab[2 * j]) % MINSTD_M;
        const uint64_t x2 = (x0 * a.powtab[2 * j + 1]) % MINSTD_M;
        unif[j] = canonical_from_draws(x1, x2);
    }
    for (int w = tid; w < nw; w += nth) mask[w] = 0u;
    __syncthreads();

    const double dpos = a.delta[2 * b], dneg = a.delta[2 * b + 1];
    int kc = 0, nacc = 0;
    for (int i = 0; i < N; ++i) {
        const double gii = diag0[i] + dcorr[i];
        const int si = sl[i];
        const double delta = si > 0 ? dpos : dneg;
        double x;
        const double p = accept_prob(gii, delta, &x);
        if (unif[i] < p) {
            const double gam = delta / x;
            for (int j = tid; j < N; j += nth) {
                T aj = G[j + (long long)i * N];
                T bj = G[i + (long long)j * N];
                for (int m = 0; m < kc; ++m) {
                    aj += X[m * N + j] * Y[m * N + i];
                    bj += X[m * N + i] * Y[m * N + j];
                }
                if (j == i) bj -= T(1);
                const T xv = T(gam * (double)aj);
                X[kc * N + j] = xv;
                Y[kc * N + j] = bj;
                dcorr[j] += (double)xv * (double)bj;
            }
            if (tid == 0) {
                sl[i] = -si;
                f[i] = (int8_t)(-si);
                mask[i >> 5] |= 1u << (i & 31);
            }
            ++kc;
            ++nacc;
            __syncthreads();
            if (kc == kmax) {
                flush_pending<T>(G, N, X, Y, diag0, dcorr, kc, kmax, tid, nth);
                kc = 0;
            }
        }
    }
    if (kc > 0) flush_pending<T>(G, N, X, Y, diag0, dcorr, kc, kmax, tid, nth);
    __syncthreads();
    uint32_t* gm = a.accmask + ((long long)b * a.L + a.t) * nw;
    for (int w = tid; w < nw; w += nth) gm[w] = mask[w];
    if (tid == 0) {
        a.rng[b] = (x0 * a.powtab[2 * N - 1]) % MINSTD_M;
        a.acc_count[b] += (unsigned long long)nacc;
    }
}

size_t metropolis_lds_bytes(int N, int krows, size_t esz)
{
    return (size_t)2 * krows * N * esz + (size_t)3 * N * sizeof(double) + (size_t)N * sizeof(int) +
           (size_t)((N + 31) / 32) * sizeof(uint32_t) + 64;
}

hipError_t prepare_kernels()
{
    hipError_t e = hipFuncSetAttribute((const void*)metropolis_kernel<double>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void*)metropolis_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               160 * 1024);
}

template <typename T>
hipError_t launch_metropolis(const MetroArgs<T>& a, int batch, hipStream_t s)
{
    const int kal = a.kmax == 1 ? 1 : ((a.kmax + 7) / 8) * 8;
    const size_t shm = metropolis_lds_bytes(a.N, kal, sizeof(T));
    const int threads = a.N >= 256 ? 256 : ((a.N + 63) / 64) * 64;
    hipLaunchKernelGGL(metropolis_kernel<T>, dim3(batch), dim3(threads), shm, s, a);
    return hipGetLastError();
}
template hipError_t launch_metropolis<double>(const MetroArgs<double>&, int, hipStream_t);
template hipError_t launch_metropolis<float>(const MetroArgs<float>&, int, hipStream_t);

// ---------------------------------------------------------------------------------------------
//  Blocked Householder QR of a row-major N x N block with Q^T applied to W-N trailing columns.
//  One 256-thread workgroup per chain.  Panel of NB columns lives in LDS; its reflectors
//  (LAPACK dlarfg convention) form the compact WY factor T; the trailing update
//  A <- (I - V T^T V^T) A runs thread-per-column (coalesced along rows of the row-major buffer).
// ---------------------------------------------------------------------------------------------
#define QR_NB 16
#define QR_MAXN 256

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void qr_kernel(double* __restrict__ bufs, long long stride, int N, int W)
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

hipError_t launch_qr(double* buf, long long stride, int N, int W, int batch, hipStream_t s)
{
    if (N > QR_MAXN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(qr_kernel, dim3(batch), dim3(256), 0, s, buf, stride, N, W);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
//  Upper triangular solve, thread-per-right-hand-side column, 16-row blocks from the bottom.
// ---------------------------------------------------------------------------------------------
#define TRSM_BR 16
__global__ __launch_bounds__(256) void trsm_kernel(double* __restrict__ bufs, long long stride, int N, int W,
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

hipError_t launch_trsm_upper(double* buf, long long stride, int N, int W, int c0, int P, int batch,
                             hipStream_t s)
{
    dim3 grid((P + 255) / 256, batch);
    hipLaunchKernelGGL(trsm_kernel, grid, dim3(256), 0, s, buf, stride, N, W, c0, P);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
//  UDV helpers
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void udv_prepare_kernel(const double* __restrict__ Ms, long long sM,
                                                          double* __restrict__ bufs, long long sbuf,
                                                          int* __restrict__ perms, int N)
{
    __shared__ double nrm[1024];
    const int b = blockIdx.x, tid = threadIdx.x;
    const double* M = Ms + (long long)b * sM;
    double* buf = bufs + (long long)b * sbuf;
    int* perm = perms + (long long)b * N;
    const int W = 2 * N;
    for (int j = tid; j < N; j += 256) {
        double s = 0.0;
        for (int r = 0; r < N; ++r) { const double v = M[r + (long long)j * N]; s += v * v; }
        nrm[j] = s;
    }
    __syncthreads();
    // rank by descending norm, ties broken by index (a stable sort)
    for (int j = tid; j < N; j += 256) {
        const double nj = nrm[j];
        int rank = 0;
        for (int i = 0; i < N; ++i) {
            const double ni = nrm[i];
            rank += (ni > nj) || (ni == nj && i < j);
        }
        perm[rank] = j;
    }
    __syncthreads();
    for (int e = tid; e < N * N; e += 256) {
        const int r = e / N, j = e % N;
        buf[(long long)r * W + j] = M[r + (long long)perm[j] * N];
        buf[(long long)r * W + N + j] = (r == j) ? 1.0 : 0.0;
    }
}

hipError_t launch_udv_prepare(const double* M, long long sM, double* buf, long long sbuf, int* perm, int N,
                              int batch, hipStream_t s)
{
    if (N > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(udv_prepare_kernel, dim3(batch), dim3(256), 0, s, M, sM, buf, sbuf, perm, N);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void udv_finish_kernel(const double* __restrict__ bufs, long long sbuf,
                                                         const int* __restrict__ perms, double* __restrict__ Us,
                                                         long long sU, double* __restrict__ Ds, long long sD,
                                                         double* __restrict__ Tts, long long sTt, int N)
{
    const int b = blockIdx.y;
    const double* buf = bufs + (long long)b * sbuf;
    const int* perm = perms + (long long)b * N;
    double* U = Us + (long long)b * sU;
    double* D = Ds + (long long)b * sD;
    double* Tt = Tts + (long long)b * sTt;
    const int W = 2 * N;
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)N * N) return;
    const int i = (int)(e / N), j = (int)(e % N);
    // U column-major: U(j, i) = Q(j, i) = (Q^T)(i, j) = buf[i][N + j]
    U[j + (long long)i * N] = buf[(long long)i * W + N + j];
    const double d = fabs(buf[(long long)i * W + i]);
    if (j == 0) D[i] = d;
    // Tt(perm[j], i) = R(i, j) / D_i for j >= i
    Tt[perm[j] + (long long)i * N] = (j >= i) ? buf[(long long)i * W + j] / d : 0.0;
}

hipError_t launch_udv_finish(const double* buf, long long sbuf, const int* perm, double* U, long long sU,
                             double* D, long long sD, double* Tt, long long sTt, int N, int batch, hipStream_t s)
{
    dim3 grid((unsigned)(((long long)N * N + 255) / 256), batch);
    hipLaunchKernelGGL(udv_finish_kernel, grid, dim3(256), 0, s, buf, sbuf, perm, U, sU, D, sD, Tt, sTt, N);
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

__global__ __launch_bounds__(256) void compare_commit_kernel(const double* __restrict__ Gnew, double* __restrict__ G64,
                                                             float* __restrict__ G32, long long sG, int N,
                                                             double* err_signed, double* err_abs)
{
    __shared__ double red[4];
    const int b = blockIdx.x;
    const double* gn = Gnew + (long long)b * sG;
    double* g64 = G64 + (long long)b * sG;
    float* g32 = G32 ? G32 + (long long)b * sG : nullptr;
    double es = -DBL_MAX, ea = 0.0;
    int bad = 0;
    for (long long e = threadIdx.x; e < (long long)N * N; e += 256) {
        const double v = gn[e];
        const double cur = g32 ? (double)g32[e] : g64[e];
        const double d = v - cur;
        es = fmax(es, d);
        ea = fmax(ea, fabs(d));
        bad |= !isfinite(v);
        g64[e] = v;
        if (g32) g32[e] = (float)v;
    }
    es = block_max(es, red);
    ea = block_max(ea, red);
    const double nanflag = block_max(bad ? 1.0 : 0.0, red);
    if (threadIdx.x == 0) {
        if (es > err_signed[b]) err_signed[b] = es;
        if (ea > err_abs[b] || nanflag > 0.0) err_abs[b] = nanflag > 0.0 ? INFINITY : fmax(err_abs[b], ea);
    }
}

hipError_t launch_compare_commit(const double* Gnew, double* G64, float* G32, long long sG, int N,
                                 double* err_signed, double* err_abs, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(compare_commit_kernel, dim3(batch), dim3(256), 0, s, Gnew, G64, G32, sG, N, err_signed,
                       err_abs);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void compare_only_kernel(const double* __restrict__ Gnew,
                                                           const double* __restrict__ Gcur, long long sG, int N,
                                                           double* err_signed)
{
    __shared__ double red[4];
    const int b = blockIdx.x;
    double es = -DBL_MAX;
    for (long long e = threadIdx.x; e < (long long)N * N; e += 256)
        es = fmax(es, Gnew[(long long)b * sG + e] - Gcur[(long long)b * sG + e]);
    es = block_max(es, red);
    if (threadIdx.x == 0 && es > err_signed[b]) err_signed[b] = es;
}

hipError_t launch_compare_only(const double* Gnew, const double* Gcur, long long sG, int N, double* err_signed,
                               int batch, hipStream_t s)
{
    hipLaunchKernelGGL(compare_only_kernel, dim3(batch), dim3(256), 0, s, Gnew, Gcur, sG, N, err_signed);
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

}  // namespace dqmc

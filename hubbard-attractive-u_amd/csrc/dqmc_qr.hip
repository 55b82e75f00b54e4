// dqmc_qr.hip -- blocked Householder QR (panel in registers, trailing update on MFMA), blocked
// partial-pivot LU and blocked back-substitution of the row-major stabilization systems, plus the
// one-kernel QR / trsm variants.
#include "dqmc_device.h"

namespace dqmc {

#define QR_NB 16      // panel width of the one-kernel QR (qr_v1_kernel)
#define QR_MAXN 256

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
#define QR_WIDE_MAXN 1024   // largest system of the blocked QR (more than 256 rows: wide panel kernel)

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
    // latency-bound: issue ahead of the update kernels that share the CU
    __builtin_amdgcn_s_setprio(3);
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

// Row-major write-back of a 256-thread panel held one row per thread (thread r: row r, x[0..QRB-1])
// through an LDS transpose in two halves of 16 columns: every store instruction then writes whole
// 128-byte row segments (4 rows x 16 columns per wave) instead of 64 scattered doubles.  Row r,
// column m of the panel goes to dst0[r * ld0 + m] as f0(r, m, x) and, when dst1 is set, to
// dst1[r * ld1 + m] as f1(r, m, x), for r < R and m < nb.  All 256 threads must call it.
template <typename F0, typename F1>
__device__ __forceinline__ void panel_rows_store(double (*tp)[QRB / 2 + 1], const double (&x)[QRB], int R, int nb,
                                                 double* __restrict__ dst0, long long ld0, F0 f0,
                                                 double* __restrict__ dst1, long long ld1, F1 f1)
{
    const int r = threadIdx.x;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h) __syncthreads();   // the first half's reads are done before its slots are reused
#pragma unroll
        for (int j = 0; j < QRB / 2; ++j) tp[r][j] = x[QRB / 2 * h + j];
        __syncthreads();
#pragma unroll 4
        for (int e = r; e < 256 * (QRB / 2); e += 256) {
            const int row = e >> 4, m = QRB / 2 * h + (e & 15);
            if (row < R && m < nb) {
                const double v = tp[row][e & 15];
                dst0[row * ld0 + m] = f0(row, m, v);
                if (dst1) dst1[row * ld1 + m] = f1(row, m, v);
            }
        }
    }
}

// Panel factorization, 256 threads (4 waves) per chain: thread r holds panel row r (rows p0 .. N-1,
// so N - p0 <= 256) with all QRB columns in registers.  Column i costs two 4-wave barriers: one
// after the partial norms (alpha = x(i,i) is published with them), one after the partial sums of
// the 32 products v . x_j (j > i) and v_k . v (k < i, the V^T V entries of T) that one recursive
// halving reduces per wave.  The column sums reach every lane through v_readlane (uniform values;
// an LDS row read at one address per entry measured 66 -> 71 us per panel).
// Every LDS slot is written between the second barrier of one column and the first of the next,
// and read between the two barriers of the column that wrote it, so single buffers suffice.
// T (H_1..H_nb = I - V T V^T, column-major QRB x QRB) comes from the recursion
// T(j,i) = -tau_i sum_{k=j}^{i-1} T(j,k) (V^T V)(k,i), thread j forming row j on its own.
__device__ __forceinline__ void qr_update_tile(double* __restrict__ bufs, long long stride, int N, int W, int b, int cb,
                                               int p0, int nb, const double* __restrict__ Vs,
                                               const double* __restrict__ Vrows, long long vr_stride,
                                               const double* __restrict__ Ts, double (*Wp)[QRB][QRB + 1],
                                               int ident = 0x7fffffff);
// With pp0 >= 0 the workgroup first applies the previous panel's reflectors (V, T of the panel at row
// pp0, row-major copy pVr) to this panel's columns, as the trailing update's tile 0 would, and then
// factors them: one launch, and one kernel boundary, fewer per panel.  `pident` is that update's
// identity column (qr_update_tile).
__global__ __launch_bounds__(256) void qr_panel_kernel(double* __restrict__ bufs, long long stride, int N, int W,
                                                       int p0, int nb, double* __restrict__ Vs,
                                                       double* __restrict__ Ts, double* __restrict__ Vrows,
                                                       long long vr_stride, int pp0, const double* __restrict__ pV,
                                                       const double* __restrict__ pT, const double* __restrict__ pVr,
                                                       int pident)
{
    // latency-bound: issue ahead of the update kernels that share the CU
    __builtin_amdgcn_s_setprio(3);
    __shared__ double nrm[4];
    __shared__ double alph;
    __shared__ double dot[4][QRB];
    __shared__ __align__(16) double VtVt[QRB][QRB];   // VtVt[i][k] = (V^T V)(k, i), k < i
    __shared__ double tau_s[QRB];
    __shared__ double tp[256][QRB / 2 + 1];            // write-back transpose
    const int r = threadIdx.x, lane = r & 63, wave = r >> 6;
    const int b = blockIdx.x;
    const int R = N - p0;
    const bool live = r < R;
    double* A = bufs + (long long)b * stride;
    double* V = Vs + (long long)b * N * QRB;
    double* Tg = Ts + (long long)b * QRB * QRB;
    static_assert(sizeof(tp) >= 4 * QRB * (QRB + 1) * sizeof(double), "update scratch");
    if (pp0 >= 0) {
        // tp is free until the write-back; this workgroup's stores of the tile are visible to it past
        // the barrier
        qr_update_tile(bufs, stride, N, W, b, 0, pp0, (N - pp0 < QRB) ? N - pp0 : QRB, pV, pVr, vr_stride, pT,
                       reinterpret_cast<double(*)[QRB][QRB + 1]>(&tp[0][0]), pident);
        __syncthreads();
    }

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
        for (int m = 0; m < QRB; ++m)
            if (m < nb) V[r + (long long)m * N] = (r == m) ? 1.0 : (r > m ? x[m] : 0.0);
    }
    // R back into the system and the row-major copy of V (rows at QRB), as whole row segments
    panel_rows_store(tp, x, R, nb, A + (long long)p0 * W + p0, W,
                     [](int row, int m, double v) { return row <= m ? v : 0.0; },
                     Vrows ? Vrows + (long long)b * vr_stride : nullptr, QRB,
                     [](int row, int m, double v) { return row == m ? 1.0 : (row > m ? v : 0.0); });
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
// The update of column tile cb of chain b by one 256-thread workgroup (all threads call it); Wp is
// 4 x QRB x (QRB + 1) doubles of LDS.  Columns from `ident` on are taken as those of the identity,
// column ident + j being e_j (the untouched right-hand side of the first panel's update), instead of
// being read.
__device__ __forceinline__ void qr_update_tile(double* __restrict__ bufs, long long stride, int N, int W, int b, int cb,
                                               int p0, int nb, const double* __restrict__ Vs,
                                               const double* __restrict__ Vrows, long long vr_stride,
                                               const double* __restrict__ Ts, double (*Wp)[QRB][QRB + 1],
                                               int ident)
{
    typedef Mfma<double>::acc_t acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane >> 4, lc = lane & 15;
    const int R = N - p0, cols = W - (p0 + nb), nrt = (R + 15) >> 4;
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
                const int a = p0 + nb + c0 + col;   // the column in the system
                at[q][ct][r] = !(row < R && c0 + col < cols) ? 0.0
                               : (a >= ident ? (p0 + row == a - ident ? 1.0 : 0.0) : A[base + (long long)row * W + col]);
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

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void qr_update_kernel(double* __restrict__ bufs, long long stride, int N,
                                                            int W, int batch, int p0, int nb, int cb0, int ncb,
                                                            const double* __restrict__ Vs,
                                                            const double* __restrict__ Vrows, long long vr_stride,
                                                            const double* __restrict__ Ts, int ident)
{
    __shared__ double Wp[4][QRB][QRB + 1];
    int b, cb;
    chain_tile(ncb, batch, b, cb);
    qr_update_tile(bufs, stride, N, W, b, cb + cb0, p0, nb, Vs, Vrows, vr_stride, Ts, Wp, ident);
}


// Panel factorization for systems of more than 256 rows (N <= 1024): NW waves per chain, thread t
// holds the RPT panel rows t, t + 64 NW, .. with all QRB columns in registers.  The column steps,
// the two barriers per column, the recursive-halving reductions and the T recursion are those of
// qr_panel_kernel; the per-wave partials are summed over the NW waves in wave order.
template <int NW, int RPT>
__global__ __launch_bounds__(64 * NW) void qr_panel_wide_kernel(double* __restrict__ bufs, long long stride, int N,
                                                                int W, int p0, int nb, double* __restrict__ Vs,
                                                                double* __restrict__ Ts)
{
    // latency-bound: issue ahead of the update kernels that share the CU
    __builtin_amdgcn_s_setprio(3);
    constexpr int NT = 64 * NW;
    __shared__ double nrm[NW];
    __shared__ double alph;
    __shared__ double dot[NW][QRB];
    __shared__ __align__(16) double VtVt[QRB][QRB];   // VtVt[i][k] = (V^T V)(k, i), k < i
    __shared__ double tau_s[QRB];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int b = blockIdx.x;
    const int R = N - p0;
    double* A = bufs + (long long)b * stride;
    double* V = Vs + (long long)b * N * QRB;
    double* Tg = Ts + (long long)b * QRB * QRB;

    double x[RPT][QRB];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = t + k * NT;
#pragma unroll
        for (int m = 0; m < QRB; ++m) x[k][m] = (r < R && m < nb) ? A[(long long)(p0 + r) * W + p0 + m] : 0.0;
    }
    if (t < QRB) tau_s[t] = 0.0;
    for (int e = t; e < QRB * QRB; e += NT) VtVt[e / QRB][e % QRB] = 0.0;

#pragma clang loop unroll(full)
    for (int i = 0; i < QRB; ++i) {
        if (i >= nb) continue;   // uniform
        double n2 = 0.0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = t + k * NT;
            if (r == i) alph = x[k][i];
            if (r < R && r > i) n2 += x[k][i] * x[k][i];
        }
        n2 = wave_sum_dpp(n2);
        if (lane == 0) nrm[wave] = n2;
        __syncthreads();
        double s2 = nrm[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) s2 += nrm[w];
        const double alpha = alph;
        double beta = alpha, tau = 0.0, scal = 0.0;
        if (s2 > 0.0) {
            beta = -copysign(sqrt(alpha * alpha + s2), alpha);
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        double v[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = t + k * NT;
            v[k] = (r == i) ? 1.0 : ((r < R && r > i) ? x[k][i] * scal : 0.0);
            if (r == i) x[k][i] = beta;
            else if (r > i) x[k][i] = v[k];
        }
        // m < i: v_m . v_i; m > i: v_i . x_m (rows r < i have v = 0, rows r >= i hold v_m(r) for m < i)
        double d[QRB];
#pragma unroll
        for (int m = 0; m < QRB; ++m) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) s += v[k] * x[k][m];
            d[m] = (m == i) ? 0.0 : s;
        }
        const double ws = wave_halve32(d, lane);
        if ((lane & 1) == 0) dot[wave][halve32_index(lane)] = ws;
        if (t == 0) tau_s[i] = tau;
        __syncthreads();
        double wsum = 0.0;
        if (lane < QRB) {
            wsum = dot[0][lane];
#pragma unroll
            for (int w = 1; w < NW; ++w) wsum += dot[w][lane];
        }
        if (t < i) VtVt[i][t] = wsum;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const double tv = tau * v[k];
#pragma unroll
            for (int m = 0; m < QRB; ++m)
                if (m > i) x[k][m] -= tv * readlane_f64(wsum, m);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = t + k * NT;
        if (r >= R) continue;
#pragma unroll
        for (int m = 0; m < QRB; ++m) {
            if (m < nb) {
                A[(long long)(p0 + r) * W + p0 + m] = (r <= m) ? x[k][m] : 0.0;
                V[r + (long long)m * N] = (r == m) ? 1.0 : (r > m ? x[k][m] : 0.0);
            }
        }
    }
    if (t < QRB) {
        const int j = t;
        double tr[QRB];
#pragma unroll
        for (int i = 0; i < QRB; ++i) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < i; ++k) s += tr[k] * VtVt[i][k];
            const double tt = (i == j) ? tau_s[i] : -tau_s[i] * s;
            tr[i] = (i < j) ? 0.0 : tt;
        }
#pragma unroll
        for (int k = 0; k < QRB; ++k) Tg[j + (long long)k * QRB] = tr[k];
    }
}

hipError_t launch_qr_blocked(double* buf, long long stride, int N, int W, int batch, double* Vs, double* Ts,
                             double* Wt, long long wt_stride, int tile, int update, int panel, hipStream_t s,
                             const LookAhead* la, bool rhs_identity)
{
    // the first panel's update reads the columns from N on as the identity (QR_UPDATE_REG only)
    if (rhs_identity && (update != QR_UPDATE_REG || N > 256 || W < 2 * N)) return hipErrorInvalidValue;
    const int ident0 = rhs_identity ? N : 0x7fffffff;
    // per chain: Vs N*QRB, Ts QRB*QRB, Wt 2 * wt_stride (Wt and Wt T; in QR_UPDATE_REG the first
    // holds the row-major copy of V, the second that of the look-ahead's parity-1 V), wt_stride >= W*QRB
    if (N > QR_WIDE_MAXN || N < 1 || W <= N) return hipErrorInvalidValue;
    if (wt_stride < (long long)W * QRB) return hipErrorInvalidValue;
    // more than 256 rows: the wide panel kernel and the three-GEMM update, in order
    if (N > 256) { update = QR_UPDATE_GEMM; la = nullptr; }
    double* Wt2 = Wt + (long long)batch * wt_stride;
    // panel at row p0 into the parity-par buffers; with pp0 >= 0 (256-thread panel only) it first applies
    // the panel at row pp0 (parity ppar) to its own columns
    auto launch_panel = [&](int p0, int par, hipStream_t st, int pp0 = -1, int ppar = 0) {
        const int nb = (N - p0 < QRB) ? N - p0 : QRB;
        double *V = par ? la->V2 : Vs, *T = par ? la->T2 : Ts, *Vr = par ? Wt2 : Wt;
        const double *pV = ppar ? la->V2 : Vs, *pT = ppar ? la->T2 : Ts, *pVr = ppar ? Wt2 : Wt;
        if (panel == QR_PANEL_1024)
            hipLaunchKernelGGL(qr_panel4_kernel, dim3(batch), dim3(1024), 0, st, buf, stride, N, W, p0, nb, V, T, Vr,
                               wt_stride);
        else
            hipLaunchKernelGGL(qr_panel_kernel, dim3(batch), dim3(256), 0, st, buf, stride, N, W, p0, nb, V, T, Vr,
                               wt_stride, pp0, pV, pT, pVr, pp0 == 0 ? ident0 : 0x7fffffff);
        return hipGetLastError();
    };
    auto launch_update = [&](int p0, int par, int cb0, int ncb, hipStream_t st) {
        const int nb = (N - p0 < QRB) ? N - p0 : QRB;
        double *V = par ? la->V2 : Vs, *T = par ? la->T2 : Ts, *Vr = par ? Wt2 : Wt;
        hipLaunchKernelGGL(qr_update_kernel, dim3(ncb * batch), dim3(256), 0, st, buf, stride, N, W, batch, p0, nb,
                           cb0, ncb, V, Vr, wt_stride, T, p0 == 0 ? ident0 : 0x7fffffff);
        return hipGetLastError();
    };
    if (la && la->s2 && update == QR_UPDATE_REG) {
        // Look-ahead: panel k + 1 (main stream) first applies update k to its own columns (block 0) and
        // then factors them, beside the rest of update k (second stream).  Panel buffers alternate by
        // parity, so panel k + 1 never overwrites what the rest of update k reads.  The 1024-thread
        // panel leaves block 0 to its own update launch.
        const int np = (N + QRB - 1) / QRB;
        const bool fused = panel != QR_PANEL_1024 && la->fuse;
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
            if (!fused && (e = launch_update(p0, k & 1, 0, 1, s)) != hipSuccess) return e;
            if (ncb > 1) {
                if ((e = hipEventRecord(la->fork, s)) != hipSuccess) return e;
                if ((e = hipStreamWaitEvent(la->s2, la->fork, 0)) != hipSuccess) return e;
            }
            // the panel is submitted first so that its workgroups take their slots before the update's
            if ((e = launch_panel(p0 + QRB, (k + 1) & 1, s, fused ? p0 : -1, k & 1)) != hipSuccess) return e;
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
        if (R > 512)
            hipLaunchKernelGGL((qr_panel_wide_kernel<8, 2>), dim3(batch), dim3(512), 0, s, buf, stride, N, W, p0, nb, Vs,
                               Ts);
        else if (R > 256)
            hipLaunchKernelGGL((qr_panel_wide_kernel<8, 1>), dim3(batch), dim3(512), 0, s, buf, stride, N, W, p0, nb, Vs,
                               Ts);
        else if (panel == QR_PANEL_1024)
            hipLaunchKernelGGL(qr_panel4_kernel, dim3(batch), dim3(1024), 0, s, buf, stride, N, W, p0, nb, Vs, Ts,
                               reg ? Wt : nullptr, wt_stride);
        else
            hipLaunchKernelGGL(qr_panel_kernel, dim3(batch), dim3(256), 0, s, buf, stride, N, W, p0, nb, Vs, Ts,
                               reg ? Wt : nullptr, wt_stride, -1, (const double*)nullptr, (const double*)nullptr,
                               (const double*)nullptr, 0x7fffffff);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (cols <= 0) continue;
        if (reg) {
            hipLaunchKernelGGL(qr_update_kernel, dim3(((cols + QRB - 1) / QRB) * batch), dim3(256), 0, s, buf,
                               stride, N, W, batch, p0, nb, 0, (cols + QRB - 1) / QRB, Vs, Wt, wt_stride, Ts,
                               p0 == 0 ? ident0 : 0x7fffffff);
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
// max of v over the 64 lanes, in every lane (max is exact, so the order of the pairings is free)
__device__ __forceinline__ double wave_max_f64(double v, int lane)
{
    v = fmax(v, dpp_f64<DPP_XOR1>(v));
    v = fmax(v, dpp_f64<DPP_XOR2>(v));
    v = fmax(v, dpp_f64<DPP_HALF_MIRROR>(v));
    v = fmax(v, dpp_f64<DPP_MIRROR>(v));
    v = fmax(v, xchg_f64(v, lane, true));
    return fmax(v, xchg_f64(v, lane, false));
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
// below eliminate with the pivot row (read from LDS at one address by every lane).
// The published slots alternate between two buffers by column parity: a slot written for column
// i + 2 is past the barrier of column i + 1, which every reader of column i reached after its reads.
__device__ __forceinline__ void lu_update_tile(double* __restrict__ bufs, long long stride, int N, int W, int b, int cb,
                                               int p0, int nb, const double* __restrict__ Ls,
                                               const double* __restrict__ Li, const int* __restrict__ Ps,
                                               double (*S)[QRB][QRB + 1]);
// With pp0 >= 0 the workgroup first applies the previous panel (L, L11^{-1} and row order of the panel
// at row pp0) to this panel's columns, as the trailing update's tile 0 would, and then factors them.
__global__ __launch_bounds__(256) void lu_panel_kernel(double* __restrict__ bufs, long long stride, int N, int W,
                                                       int p0, int nb, double* __restrict__ Ls,
                                                       double* __restrict__ Li, int* __restrict__ Ps,
                                                       uint32_t* __restrict__ health, int pp0,
                                                       const double* __restrict__ pLs, const double* __restrict__ pLi,
                                                       const int* __restrict__ pPs)
{
    // latency-bound: issue ahead of the update kernels that share the CU
    __builtin_amdgcn_s_setprio(3);
    __shared__ double redv[2][4];
    __shared__ int redi[2][4];
    __shared__ __align__(16) double cand[2][4][QRB];
    __shared__ int cando[2][4];
    __shared__ __align__(16) double irow[2][QRB];
    __shared__ int irowo[2];
    __shared__ double L11[QRB][QRB + 1];
    __shared__ double tp[256][QRB / 2 + 1];            // write-back transpose
    const int r = threadIdx.x, lane = r & 63, wave = r >> 6;
    const int b = blockIdx.x;
    const int R = N - p0;
    const bool live = r < R;
    double* A = bufs + (long long)b * stride;
    static_assert(sizeof(tp) >= 2 * QRB * (QRB + 1) * sizeof(double), "update scratch");
    if (pp0 >= 0) {
        // tp is free until the write-back; this workgroup's stores of the tile are visible to it past
        // the barrier
        lu_update_tile(bufs, stride, N, W, b, 0, pp0, (N - pp0 < QRB) ? N - pp0 : QRB, pLs, pLi, pPs,
                       reinterpret_cast<double(*)[QRB][QRB + 1]>(&tp[0][0]));
        __syncthreads();
    }

    double x[QRB];
#pragma unroll
    for (int m = 0; m < QRB; ++m) x[m] = (live && m < nb) ? A[(long long)(p0 + r) * W + p0 + m] : 0.0;
    int org = r;
    bool zero_pivot = false;

#pragma clang loop unroll(full)
    for (int i = 0; i < QRB; ++i) {
        if (i >= nb) continue;   // uniform
        const int par = i & 1;
        // the wave's candidate: the largest |x(r, i)|, ties to the lowest row (lowest lane of the
        // ballot); a max reduction and a ballot are a shorter dependent chain than a (value, index)
        // arg-max
        const double cl = (live && r >= i) ? fabs(x[i]) : -1.0;
        const double cv = wave_max_f64(cl, lane);
        const unsigned long long hit = __ballot(cl == cv);
        const int lmin = hit ? __builtin_ctzll(hit) : 0;
        const int ci = wave * 64 + lmin;
        if (lane == lmin) {   // this lane holds its wave's candidate row
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
        // the pivot row, read by every lane at one address (LDS broadcast): 16 loads of two entries
        // instead of 62 v_readlane of a lane-distributed copy (69.0 -> 66.1 us per panel alone)
        const double* prow = cand[par][bw];
        if (r == i && bi != i) {
#pragma unroll
            for (int m = 0; m < QRB; ++m) x[m] = prow[m];
            org = cando[par][bw];
        } else if (r == bi && bi != i) {
#pragma unroll
            for (int m = 0; m < QRB; ++m) x[m] = irow[par][m];
            org = irowo[par];
        }
        const double piv = prow[i];
        zero_pivot |= !(piv != 0.0);
        const double l = (live && r > i) ? x[i] / piv : 0.0;
        if (r > i) x[i] = l;
#pragma unroll
        for (int m = (i + 1) & ~1; m < QRB; m += 2) {
            const double2 pm = *reinterpret_cast<const double2*>(prow + m);
            if (m > i) x[m] -= l * pm.x;
            x[m + 1] -= l * pm.y;
        }
    }
    if (r == 0 && zero_pivot) atomicOr(health + b, HEALTH_ZERO_PIVOT);
    if (live) {
#pragma unroll
        for (int m = 0; m < QRB; ++m)
            if (m < nb) Ls[(long long)b * N * QRB + r + (long long)m * N] = (r > m) ? x[m] : (r == m ? 1.0 : 0.0);
        Ps[(long long)b * N + r] = org;
    }
    // L and U back into the system, as whole row segments
    panel_rows_store(tp, x, R, nb, A + (long long)p0 * W + p0, W, [](int, int, double v) { return v; },
                     (double*)nullptr, 0, [](int, int, double v) { return v; });
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
// and 1 hold the QRB panel rows A12: through LDS the four waves form U12 = L11^{-1} A12 (one 16 x 16
// tile each), which every wave then reads as B-fragments for A22 -= L21 U12 on its own row tiles.  The
// tile is read and written once.
// Four waves per SIMD (at most 128 registers): 145 instead of 164 us for the first panel's update at
// 256 chains.
// The update of column tile cb of chain b by one 256-thread workgroup (all threads call it); S is
// 2 x QRB x (QRB + 1) doubles of LDS ([0]: A12, [1]: U12).
__device__ __forceinline__ void lu_update_tile(double* __restrict__ bufs, long long stride, int N, int W, int b, int cb,
                                               int p0, int nb, const double* __restrict__ Ls,
                                               const double* __restrict__ Li, const int* __restrict__ Ps,
                                               double (*S)[QRB][QRB + 1])
{
    typedef Mfma<double>::acc_t acc_t;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lr = lane >> 4, lc = lane & 15;
    const int R = N - p0, cols = W - (p0 + nb), nrt = (R + 15) >> 4;
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
    {
        // U12 tile (wave / 2, wave % 2) by each wave
        const int mt = wave >> 1, ct = wave & 1;
        acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < QRB; k0 += 4)
            acc = Mfma<double>::mma(Lv[(16 * mt + lc) + (long long)(k0 + lr) * QRB], S[0][k0 + lr][16 * ct + lc], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) S[1][16 * mt + 4 * r + lr][16 * ct + lc] = acc[r];
    }
    __syncthreads();
    if (wave < 2) {
        // the U12 rows of this wave's first row tile
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) at[0][ct][r] = S[1][16 * wave + 4 * r + lr][16 * ct + lc];
    }
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

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void lu_update_kernel(double* __restrict__ bufs, long long stride, int N, int W,
                                                        int batch, int p0, int nb, int cb0, int ncb,
                                                        const double* __restrict__ Ls, const double* __restrict__ Li,
                                                        const int* __restrict__ Ps)
{
    __shared__ double S[2][QRB][QRB + 1];
    int b, cb;
    chain_tile(ncb, batch, b, cb);
    lu_update_tile(bufs, stride, N, W, b, cb + cb0, p0, nb, Ls, Li, Ps, S);
}


hipError_t launch_lu_blocked(double* buf, long long stride, int N, int W, int batch, double* Ls, double* Li, int* Ps,
                             uint32_t* health, hipStream_t s, const LookAhead* la)
{
    if (N > 256 || N < 1 || W <= N) return hipErrorInvalidValue;
    const bool ahead = la && la->s2;
    // panel at row p0 into the parity-par buffers; with pp0 >= 0 it first applies the panel at row pp0
    // (parity ppar) to its own columns
    auto launch_panel = [&](int p0, int par, hipStream_t st, int pp0 = -1, int ppar = 0) {
        const int nb = (N - p0 < QRB) ? N - p0 : QRB;
        hipLaunchKernelGGL(lu_panel_kernel, dim3(batch), dim3(256), 0, st, buf, stride, N, W, p0, nb,
                           par ? la->V2 : Ls, par ? la->T2 : Li, par ? la->P2 : Ps, health, pp0,
                           (const double*)(ppar ? la->V2 : Ls), (const double*)(ppar ? la->T2 : Li),
                           (const int*)(ppar ? la->P2 : Ps));
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
        // look-ahead, as in launch_qr_blocked: panel k + 1 applies update k to its own columns and
        // factors them beside the rest of the update (panel buffers and row orders alternate by parity)
        if (!la->fuse && (e = launch_update(p0, par, 0, 1, s)) != hipSuccess) return e;
        if (ncb > 1) {
            if ((e = hipEventRecord(la->fork, s)) != hipSuccess) return e;
            if ((e = hipStreamWaitEvent(la->s2, la->fork, 0)) != hipSuccess) return e;
        }
        if ((e = launch_panel(p0 + QRB, npar, s, la->fuse ? p0 : -1, par)) != hipSuccess) return e;
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
//  columns [c0, c0+P), solved in place), TB-row diagonal blocks, recursively from the bottom half:
//    trsm_diag_kernel  X_k = R_kk^{-1} Z_k, one thread per right-hand side, R_kk in LDS;
//    GEMM (MFMA)       Z_lo -= R_{lo, hi} X_hi, as Z^T_{:, lo} -= X_hi^T R_{lo, hi}^T.
//  The solve is a chain of dependent launches (2 N / TB - 1 of them), so the block size trades the
//  diagonal kernels' serial work (TB divisions and TB^2 / 2 FMAs per thread) against launches.
// ---------------------------------------------------------------------------------------------
template <int TB>
__global__ __launch_bounds__(256) void trsm_diag_kernel(double* __restrict__ bufs, long long stride, int W, int c0,
                                                        int P, int k0, int nbk)
{
    // R_kk transposed (Rt[q][m] = R(m, q)), so that the column a step reads is contiguous: two
    // entries per LDS load, every lane at the same address
    __shared__ __align__(16) double Rt[TB][TB + 2];
    double* A = bufs + (long long)blockIdx.y * stride;
    for (int e = threadIdx.x; e < TB * TB; e += 256) {
        const int m = e / TB, q = e % TB;
        Rt[q][m] = (m < nbk && q < nbk) ? A[(long long)(k0 + m) * W + k0 + q] : 0.0;
    }
    __syncthreads();
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= P) return;
    double* Z = A + c0 + j;
    double x[TB];
#pragma unroll
    for (int m = 0; m < TB; ++m) x[m] = (m < nbk) ? Z[(long long)(k0 + m) * W] : 0.0;
    // Right-looking substitution: once x_m is final, every row above it takes its term at once, so each
    // step is one division and independent FMAs instead of a dependent chain of up to TB - 1 FMAs.
#pragma unroll
    for (int m = TB - 1; m >= 0; --m) {
        if (m < nbk) {
            const double xm = x[m] / Rt[m][m];
            x[m] = xm;
#pragma unroll
            for (int p = 0; p < m; ++p) x[p] -= Rt[m][p] * xm;
        }
    }
#pragma unroll
    for (int m = 0; m < TB; ++m)
        if (m < nbk) Z[(long long)(k0 + m) * W] = x[m];
}

// Blocks [lo, hi) of TB rows: the lower half of the range is updated with the solved upper half in
// one GEMM, Z_{lo:mid} -= R_{lo:mid, mid:hi} X_{mid:hi}, so each level of the recursion passes over half of
// Z once (log2(N / TB) passes, against N / TB - 1 rank-TB updates of growing height bottom-up).
template <int TB>
static hipError_t trsm_range(double* buf, long long stride, int N, int W, int c0, int P, int batch, int tile, int lo,
                             int hi, hipStream_t s)
{
    if (hi - lo == 1) {
        const int k0 = lo * TB, nbk = (N - k0 < TB) ? N - k0 : TB;
        hipLaunchKernelGGL(trsm_diag_kernel<TB>, dim3((P + 255) / 256, batch), dim3(256), 0, s, buf, stride, W, c0, P,
                           k0, nbk);
        return hipGetLastError();
    }
    const int mid = (lo + hi) / 2;
    hipError_t e = trsm_range<TB>(buf, stride, N, W, c0, P, batch, tile, mid, hi, s);
    if (e != hipSuccess) return e;
    const int r0 = lo * TB, s0 = mid * TB, s1 = (hi * TB < N) ? hi * TB : N;
    GemmArgs<double> g;
    g.M = P; g.Nc = s0 - r0; g.K = s1 - s0;
    g.A = buf + c0 + (long long)s0 * W; g.sA = stride; g.lda = W; g.tA = 0;   // X_{mid:hi}^T
    g.B = buf + (long long)r0 * W + s0; g.sB = stride; g.ldb = W; g.tB = 0;   // R_{lo:mid, mid:hi}^T
    g.C = buf + c0 + (long long)r0 * W; g.sC = stride; g.ldc = W; g.tC = 0;   // Z_{lo:mid}^T
    g.alpha = -1.0; g.beta = 1.0; g.batch = batch;
    if ((e = launch_gemm<double>(g, tile, s)) != hipSuccess) return e;
    return trsm_range<TB>(buf, stride, N, W, c0, P, batch, tile, lo, mid, s);
}

hipError_t launch_trsm_blocked(double* buf, long long stride, int N, int W, int c0, int P, int batch, int tile,
                               hipStream_t s, int block)
{
    if (N < 1 || P < 1 || c0 < N || c0 + P > W) return hipErrorInvalidValue;
    switch (block) {
    case 32: return trsm_range<32>(buf, stride, N, W, c0, P, batch, tile, 0, (N + 31) / 32, s);
    case 64: return trsm_range<64>(buf, stride, N, W, c0, P, batch, tile, 0, (N + 63) / 64, s);
    case 128: return trsm_range<128>(buf, stride, N, W, c0, P, batch, tile, 0, (N + 127) / 128, s);
    default: return hipErrorInvalidValue;
    }
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


}  // namespace dqmc

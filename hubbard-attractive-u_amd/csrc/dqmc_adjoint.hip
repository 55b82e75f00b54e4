// dqmc_adjoint.hip -- gfx950 kernels of the parameter adjoint at fixed Ising fields.
//
// For B_t = diag(e^{alpha s_t}) exp(dt (nnt A + mu)) and G(t) = (1 + B_t..B_1 B_L..B_{t+1})^{-1},
// a perturbation B_tau -> (1 + eps D_tau) B_tau changes
//     log det(1 + B_L..B_1)  by  eps tr[(1 - G(tau)) D_tau]
//     G(t)                   by  eps G(t,tau) D_tau G(tau,t)
// with the time-displaced G(t,tau), G(tau,t) (at t = tau: G(tau) on the left, -(1 - G(tau)) on
// the right).  For the three parameters D_tau is
//     nnt: dt V_tau A V_tau^{-1},   mu: dt I,   u: alpha'(u) diag(s_tau).
// A cotangent gbar_t(i) on G(t)_ii therefore contracts to sum_tau tr[D_tau Lambda_tau] with
// Lambda_tau = sum_t G(tau,t) diag(gbar_t) G(t,tau).  The engine obtains G(tau+t', tau) and
// G(tau, tau+t') as the dynamic G(t',0), G(0,t') of the fields rolled by tau slices (the
// B-chain is cyclic; the anti-periodic signs cancel in the product), so one dynamic pass per tau
// feeds the contraction below.
#include <hip/hip_runtime.h>

#include "dqmc_kernels.h"

namespace dqmc {

namespace {

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the 256 threads of a block; every thread receives the total.
__device__ __forceinline__ double block_sum_d(double v, double* red)
{
    v = wave_sum_d(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    const double r = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return r;
}

}  // namespace

// grid (L + 1, batch), 256 threads.  Block (t', b) with t' < L contracts the pair
// X = G(tau, tau+t') (vecgf0t[t']) and Y = G(tau+t', tau) (vecgft0[t']) of chain b with the
// cotangent of G(tau+t')_ii; block t' = L contracts the log-det cotangent with 1 - G(tau)
// (vecgftt[0] of the rolled pass).  partial[b][t'][0..2] = contributions to d/d(nnt, mu, u).
__global__ __launch_bounds__(256) void adjoint_contract_kernel(AdjointArgs a)
{
    __shared__ double red[4];
    const int tp = blockIdx.x, b = blockIdx.y, N = a.N;
    const long long nn = (long long)N * N;
    const int8_t* s = a.fields + (long long)b * a.L * N;   // rolled slice 0 = original slice tau
    const double ep = exp(2.0 * a.alpha), em = exp(-2.0 * a.alpha);
    double rm = 0.0, ru = 0.0, rt = 0.0;
    if (tp < a.L) {
        const double* X = a.gf0t + (long long)b * a.sv + (long long)tp * nn;
        const double* Y = a.gft0 + (long long)b * a.sv + (long long)tp * nn;
        const double* w = a.gbar_diag + ((long long)b * a.L + (a.tau + tp) % a.L) * N;
        for (int i = 0; i < N; ++i) {
            const double wi = w[i];
            if (wi == 0.0) continue;                 // uniform over the block
            const double* Xi = X + (long long)i * N;  // X(:, i)
            for (int r = threadIdx.x; r < N; r += 256) {
                const double y = Y[i + (long long)r * N];   // Y(i, r)
                const double xr = Xi[r];
                const int sr = s[r];
                double h = 0.0;
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const int q = a.nb[4 * r + d];
                    const int sq = s[q];
                    const double f = (sr == sq) ? 1.0 : (sr > sq ? ep : em);   // e^{alpha (s_r - s_q)}
                    h += f * Xi[q];
                }
                const double wy = wi * y;
                rm += wy * xr;
                ru += wy * sr * xr;
                rt += wy * h;
            }
        }
    } else {
        const double g = a.gbar_logdet ? a.gbar_logdet[b] : 0.0;
        const double* G = a.gftt + (long long)b * a.sv;
        if (g != 0.0) {
            for (int r = threadIdx.x; r < N; r += 256) {
                const double om = 1.0 - G[r + (long long)r * N];
                const int sr = s[r];
                double h = 0.0;
#pragma unroll
                for (int d = 0; d < 4; ++d) {
                    const int q = a.nb[4 * r + d];
                    const int sq = s[q];
                    const double f = (sr == sq) ? 1.0 : (sr > sq ? ep : em);
                    h += f * ((q == r ? 1.0 : 0.0) - G[q + (long long)r * N]);   // (1 - G)(q, r)
                }
                rm += g * om;
                ru += g * sr * om;
                rt += g * h;
            }
        }
    }
    rm = block_sum_d(rm, red);
    ru = block_sum_d(ru, red);
    rt = block_sum_d(rt, red);
    if (threadIdx.x == 0) {
        double* p = a.partial + ((long long)b * (a.L + 1) + tp) * 3;
        p[0] = a.dt * rt;
        p[1] = a.dt * rm;
        p[2] = a.dalpha * ru;
    }
}

// grad[b][k] += sum_{t'} partial[b][t'][k], in a fixed order (one thread per (b, k)).
__global__ void adjoint_reduce_kernel(const double* __restrict__ partial, double* __restrict__ grad, int L1, int batch)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= batch * 3) return;
    const int b = e / 3, k = e % 3;
    double s = 0.0;
    for (int t = 0; t < L1; ++t) s += partial[((long long)b * L1 + t) * 3 + k];
    grad[e] += s;
}

hipError_t launch_adjoint_contract(const AdjointArgs& a, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(adjoint_contract_kernel, dim3(a.L + 1, batch), dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int n = batch * 3;
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a.partial, a.grad, a.L + 1,
                       batch);
    return hipGetLastError();
}

// dst[b][k][i] = src[b][(k + shift) % L][i]
__global__ void roll_fields_kernel(const int8_t* __restrict__ src, int8_t* __restrict__ dst, int L, int N, int shift,
                                   long long total)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const long long per = (long long)L * N;
    const long long b = e / per, r = e % per;
    const int k = (int)(r / N), i = (int)(r % N);
    dst[e] = src[b * per + (long long)((k + shift) % L) * N + i];
}

hipError_t launch_roll_fields(const int8_t* src, int8_t* dst, int L, int N, int shift, int batch, hipStream_t s)
{
    const long long total = (long long)batch * L * N;
    hipLaunchKernelGGL(roll_fields_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, src, dst, L, N,
                       shift % L, total);
    return hipGetLastError();
}

// log|det(1 + u diag(d) v^T)| = sum_i log max(1, d_i) + sum_i log|R_ii| for the QR factor R of
// the system built by launch_build_00bb (row-major, row stride W).  One block per chain.
__global__ __launch_bounds__(256) void logdet_00bb_kernel(const double* __restrict__ bufs, long long sbuf, int W,
                                                          const double* __restrict__ ds, long long sd, int N,
                                                          double* __restrict__ out)
{
    __shared__ double red[4];
    const int b = blockIdx.x;
    double v = 0.0;
    for (int i = threadIdx.x; i < N; i += 256)
        v += log(fmax(1.0, ds[(long long)b * sd + i])) + log(fabs(bufs[(long long)b * sbuf + (long long)i * W + i]));
    v = block_sum_d(v, red);
    if (threadIdx.x == 0) out[b] = v;
}

hipError_t launch_logdet_00bb(const double* buf, long long sbuf, int W, const double* d, long long sd, int N,
                              double* out, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(logdet_00bb_kernel, dim3(batch), dim3(256), 0, s, buf, sbuf, W, d, sd, N, out);
    return hipGetLastError();
}

// out[b][t][i] = G_b(t)(i, i) from the stored equal-time slices.
__global__ void gf_diag_kernel(const double* __restrict__ vec, long long sv, int L, int N, double* __restrict__ out,
                               long long total)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const long long per = (long long)L * N;
    const long long b = e / per, r = e % per;
    const int t = (int)(r / N), i = (int)(r % N);
    out[e] = vec[b * sv + (long long)t * N * N + i + (long long)i * N];
}

hipError_t launch_gf_diag(const double* vec, long long sv, int L, int N, double* out, int batch, hipStream_t s)
{
    const long long total = (long long)batch * L * N;
    hipLaunchKernelGGL(gf_diag_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, vec, sv, L, N, out,
                       total);
    return hipGetLastError();
}

}  // namespace dqmc

// dqmc_measure.hip -- observables of the sweep path measured on the device, from the stored
// Green's functions of every slice (include/observable_methods.hpp:39-195), with per-chain bin
// caches (include/observable.hpp:160-207, src/measure_handle.cpp:68-104).
//
// One workgroup per (slice, chain) reduces its N x N matrices in a fixed order and writes one
// partial result per (chain, slice, quantity); a second launch folds the slices in order into the
// chain's cache.  No atomics, so every run gives the same bits.
#include "dqmc_kernels.h"

namespace dqmc {

__device__ __forceinline__ double wg_sum256(double v, double* red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    return ((red[0] + red[1]) + red[2]) + red[3];
}

// Equal-time quantities of slice t (both spin species share G):
//   part[0] = tr G,  part[1] = sum_i (1 - G_ii)^2,  part[2] = sum_ij (delta_ij - G_ij)^2 + G_ij^2
// (observable_methods.hpp:39-72 filling and double occupancy, :116-147 s-wave pairing, where
// sum_ij (1 - G)(j,i)^2 runs over the same entries as sum_ij (1 - G)(i,j)^2).
__global__ __launch_bounds__(256) void meas_eq_kernel(const double* __restrict__ vec, long long sv, int L, int N,
                                                      double* __restrict__ part)
{
    __shared__ double red[4];
    const int t = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    const double* G = vec + (long long)b * sv + (long long)t * N * N;
    double tr = 0.0, dd = 0.0, sw = 0.0;
    for (int e = tid; e < N * N; e += 256) {
        const int i = e % N, j = e / N;
        const double g = G[e];
        const double c = (i == j ? 1.0 : 0.0) - g;
        sw += c * c + g * g;
        if (i == j) {
            tr += g;
            dd += c * c;
        }
    }
    tr = wg_sum256(tr, red);
    dd = wg_sum256(dd, red);
    sw = wg_sum256(sw, red);
    if (tid == 0) {
        double* p = part + ((long long)b * L + t) * 3;
        p[0] = tr; p[1] = dd; p[2] = sw;
    }
}

// cache[b][0..2] += filling, double occupancy, s-wave pairing of this sample (slices in order).
__global__ void meas_eq_fold_kernel(const double* __restrict__ part, int L, int N, double* __restrict__ cache,
                                    int cstride, int nb)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double fill = 0.0, dd = 0.0, sw = 0.0;
    for (int t = 0; t < L; ++t) {
        const double* p = part + ((long long)b * L + t) * 3;
        fill += 2.0 - (p[0] + p[0]) / N;
        dd += p[1];
        sw += p[2];
    }
    double* c = cache + (long long)b * cstride;
    c[0] += fill / L;
    c[1] += dd / ((double)N * L);
    c[2] += sw / ((double)N * L);
}

// Dynamic quantities of slice t (observable_methods.hpp:84-114, :149-195):
//   G(k, t)     += (1/N) sum_ij cos(k . (r_j - r_i)) G(t,0)(j, i)
//               =  (1/N) sum_d cos(k . r_d) S(d),  S(d) = sum_i G(t,0)(i + d, i)
//   LDOS(t)     += tr G(t,0) / N
//   chi_perp(t) += -(1/N) sum_i G(0,t)_ii G(t,0)_ii
// The displacement sums S(d) take one pass over G(t,0) (thread d walks the columns, every column
// read by the workgroup as one permuted segment); the Fourier sums then run from LDS against the
// cos(k . r_d) table.  cache layout per chain: [0..2] equal-time scalars, then G(k,t) [nk][L],
// LDOS [L], chi_perp [L].
__global__ __launch_bounds__(256) void meas_dyn_kernel(const double* __restrict__ gt0, const double* __restrict__ g0t,
                                                       long long sv, int L, int nl, const double* __restrict__ costab,
                                                       int nk, double* __restrict__ cache, int cstride, int flags)
{
    __shared__ double S[1024];
    __shared__ double red[4];
    const int t = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
    const int N = nl * nl;
    const double* G = gt0 + (long long)b * sv + (long long)t * N * N;
    const double* H = g0t + (long long)b * sv + (long long)t * N * N;
    double* c = cache + (long long)b * cstride;
    if (flags & 1) {
        for (int d = tid; d < N; d += 256) {
            const int dx = d % nl, dy = d / nl;
            double s = 0.0;
            for (int i = 0; i < N; ++i) {
                const int x = (i % nl + dx) % nl, y = (i / nl + dy) % nl;
                s += G[(x + nl * y) + (long long)i * N];
            }
            S[d] = s;
        }
        __syncthreads();
        for (int k = tid; k < nk; k += 256) {
            const double* ct = costab + (long long)k * N;
            double a = 0.0;
            for (int d = 0; d < N; ++d) a += ct[d] * S[d];
            c[3 + k * L + t] += a / N;
        }
    }
    if (flags & 6) {
        double tr = 0.0, sp = 0.0;
        for (int i = tid; i < N; i += 256) {
            const double gi = G[i + (long long)i * N], hi = H[i + (long long)i * N];
            tr += gi;
            sp -= hi * gi + hi * gi;
        }
        tr = wg_sum256(tr, red);
        sp = wg_sum256(sp, red);
        if (tid == 0) {
            if (flags & 2) c[3 + nk * L + t] += tr / N;
            if (flags & 4) c[3 + nk * L + L + t] += sp / (2 * N);
        }
    }
}

// data[b][bin][...] = cache[b][...] / count, then cache = 0 (normalize_cache, push_cache_to_data and
// clear_cache of observable.hpp:160-207); eq and dyn counts apply to the two parts of the cache.
__global__ void meas_push_kernel(double* __restrict__ cache, int cstride, double* __restrict__ data,
                                 long long dstride, int bin, int neq, long long eq_count, long long dyn_count)
{
    const int b = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= cstride) return;
    double* c = cache + (long long)b * cstride;
    const long long n = e < neq ? eq_count : dyn_count;
    data[(long long)b * dstride + (long long)bin * cstride + e] = n > 0 ? c[e] / (double)n : 0.0;
    c[e] = 0.0;
}

hipError_t launch_measure_equaltime(const double* vec, long long sv, int L, int N, double* part, double* cache,
                                    int cstride, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(meas_eq_kernel, dim3(L, batch), dim3(256), 0, s, vec, sv, L, N, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(meas_eq_fold_kernel, dim3((batch + 63) / 64), dim3(64), 0, s, part, L, N, cache, cstride, batch);
    return hipGetLastError();
}

hipError_t launch_measure_dynamic(const double* gt0, const double* g0t, long long sv, int L, int nl,
                                  const double* costab, int nk, double* cache, int cstride, int flags, int batch,
                                  hipStream_t s)
{
    if (nl * nl > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(meas_dyn_kernel, dim3(L, batch), dim3(256), 0, s, gt0, g0t, sv, L, nl, costab, nk, cache,
                       cstride, flags);
    return hipGetLastError();
}

hipError_t launch_measure_push(double* cache, int cstride, double* data, long long dstride, int bin, int neq,
                               long long eq_count, long long dyn_count, int batch, hipStream_t s)
{
    hipLaunchKernelGGL(meas_push_kernel, dim3((cstride + 255) / 256, batch), dim3(256), 0, s, cache, cstride, data,
                       dstride, bin, neq, eq_count, dyn_count);
    return hipGetLastError();
}

}  // namespace dqmc

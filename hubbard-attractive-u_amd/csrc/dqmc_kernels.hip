// dqmc_kernels.hip -- small kernels of the DQMC sweep engine: UDV pre-pivot and finish, system
// builders, wrap-error reductions, the single-site update and elementwise helpers.  The large
// kernels live in dqmc_gemm.hip (GEMM, Kronecker products), dqmc_metropolis.hip and dqmc_qr.hip
// (QR, LU, back-substitution).
#include "dqmc_device.h"

namespace dqmc {

// ---------------------------------------------------------------------------------------------
//  UDV helpers
// ---------------------------------------------------------------------------------------------
// Column norms of M (column-major, ld N) and the descending-norm order (ties to the lower index,
// a stable sort): one workgroup per chain, wave w reduces the column groups 8 w .. 8 w + 7,
// 8 w + 32 .., with its lanes reading consecutive rows (lane l sums rows l, l + 64, .. in order, then
// wave_sum_dpp).  With norms_in the squared norms are read from there (N per chain) instead.
__global__ __launch_bounds__(256) void udv_rank_kernel(const double* __restrict__ Ms, long long sM,
                                                       int* __restrict__ perms, int N,
                                                       const double* __restrict__ norms_in)
{
    __shared__ double nrm[1024];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double* M = Ms + (long long)b * sM;
    int* perm = perms + (long long)b * N;
    if (norms_in)
        for (int j = tid; j < N; j += 256) nrm[j] = norms_in[(long long)b * N + j];
    // 8 columns per step, so that 8 x ceil(N / 64) loads per lane are in flight together
    for (int j0 = wave * 8; !norms_in && j0 < N; j0 += 32) {
        double s[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) s[c] = 0.0;
        for (int r = lane; r < N; r += 64) {
            double v[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) v[c] = (j0 + c < N) ? M[r + (long long)(j0 + c) * N] : 0.0;
#pragma unroll
            for (int c = 0; c < 8; ++c) s[c] = fma(v[c], v[c], s[c]);
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

// buf[r][j] = M(r, perm[j]) and, with `identity`, buf[r][N + j] = delta_rj, through a 32 x 32 LDS tile so that the
// column reads of M and the row writes of buf are both contiguous; grid (tiles, chains).
__global__ __launch_bounds__(256) void udv_copy_kernel(const double* __restrict__ Ms, long long sM,
                                                       double* __restrict__ bufs, long long sbuf,
                                                       const int* __restrict__ perms, int N, int W, int identity)
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
            if (identity) buf[(long long)r * W + N + j] = (r == j) ? 1.0 : 0.0;
        }
    }
}

hipError_t launch_udv_prepare(const double* M, long long sM, double* buf, long long sbuf, int W, int* perm, int N,
                              int batch, hipStream_t s, const double* norms, bool identity)
{
    if (N > 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(udv_rank_kernel, dim3(batch), dim3(256), 0, s, M, sM, perm, N, norms);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nt = (N + 31) / 32;
    hipLaunchKernelGGL(udv_copy_kernel, dim3(nt * nt, batch), dim3(256), 0, s, M, sM, buf, sbuf, perm, N, W,
                       identity ? 1 : 0);
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

// out = the N x N block at column c0 of the row-major system, column-major (transpose = 0: out(j, r) =
// buf[r][c0 + j]) or as is (transpose = 1: out(r, j) = buf[r][c0 + j]); 32 x 32 tiles through LDS so that
// the reads and the writes are both contiguous; grid (tiles, chains).
__global__ __launch_bounds__(256) void extract_kernel(const double* __restrict__ bufs, long long sbuf, int W,
                                                      int c0, double* __restrict__ outs, long long so, int N,
                                                      int transpose)
{
    __shared__ double t[32][33];
    const int b = blockIdx.y, nt = (N + 31) / 32;
    const int r0 = (blockIdx.x % nt) * 32, j0 = (blockIdx.x / nt) * 32;
    const double* buf = bufs + (long long)b * sbuf;
    double* out = outs + (long long)b * so;
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int rr = e >> 5, jj = e & 31;   // consecutive threads read consecutive columns of a row
        const int r = r0 + rr, j = j0 + jj;
        const double v = (r < N && j < N) ? buf[(long long)r * W + c0 + j] : 0.0;
        if (!transpose) {
            if (r < N && j < N) out[j + (long long)r * N] = v;   // already contiguous in j
        } else {
            t[jj][rr] = v;
        }
    }
    if (!transpose) return;   // uniform
    __syncthreads();
    for (int e = threadIdx.x; e < 1024; e += 256) {
        const int jj = e >> 5, rr = e & 31;   // consecutive threads write consecutive rows of a column
        const int r = r0 + rr, j = j0 + jj;
        if (r < N && j < N) out[r + (long long)j * N] = t[jj][rr];
    }
}

hipError_t launch_extract(const double* buf, long long sbuf, int W, int c0, double* out, long long so, int N,
                          int transpose, int batch, hipStream_t s)
{
    const int nt = (N + 31) / 32;
    hipLaunchKernelGGL(extract_kernel, dim3(nt * nt, batch), dim3(256), 0, s, buf, sbuf, W, c0, out, so, N, transpose);
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
#define CMP_PARTS CMP_PART_SLOTS
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

hipError_t launch_compare_fold(const double* part, int np, double* err_signed, double* err_abs, const double* bound,
                               uint32_t* health, int batch, hipStream_t s)
{
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

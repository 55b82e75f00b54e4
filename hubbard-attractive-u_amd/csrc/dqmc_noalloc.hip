// dqmc_noalloc.hip -- B-operations on caller-owned device matrices without any device allocation.
//
// dqmc_apply_bop_noalloc is the allocation-free variant of dqmc_apply_bop for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops).  The Kronecker kernels write
// out of place, so the product goes through a staging area; that area and the constant factors
// live in `scratch`, a file-scope device array that each GPU holds until the process ends.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/dqmc_amd.h"
#include "dqmc_kernels.h"

using namespace dqmc;

namespace {

// Layout of the scratch array (doubles): the 16 x 16 factors E = exp(dt nnt A1) and
// F = exp(-dt nnt A1), the diagonal tables of B and B^{-1} for s = +1, -1, then the staging area
// of n_chains N x N products at kStageOff (a 256-byte boundary).
constexpr long long kScratchBytes = DQMC_NOALLOC_SCRATCH_BYTES;
constexpr long long kE = 0, kF = 256, kTab = 512, kITab = 514, kStageOff = 1024;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) double scratch[kScratchBytes / 8];

namespace {

std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each device

struct Block256 {
    double v[256];
};

// Constants travel as a kernel argument, so no host-to-device copy is needed.
__global__ void store_block_kernel(Block256 b, double* dst, int n)
{
    const int i = threadIdx.x;
    if (i < n) dst[i] = b.v[i];
}

int scratch_for_current_device(double** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return DQMC_ENODEV;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        double* p = nullptr;
        if (hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return DQMC_EHIP;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return DQMC_OK;
}

}  // namespace

extern "C" int dqmc_apply_bop_noalloc(const dqmc_params_c* p, int n_chains, int op, const int8_t* d_fields,
                                      double* d_inout, void* stream)
{
    if (!p || !d_fields || !d_inout || n_chains < 1 || op < 0 || op > 4)
        return free_error(DQMC_EINVAL, "dqmc_apply_bop_noalloc: null pointer, n_chains < 1 or op outside 0..4");
    if (p->nl < 1 || p->nl > 16 || p->nt < 1 || p->u < 0.0 || !(p->beta >= 0.0))
        return free_error(DQMC_EINVAL, "dqmc_apply_bop_noalloc: need 1 <= nl <= 16, nt >= 1, u >= 0, beta >= 0");
    const int nl = p->nl, N = nl * nl;
    const long long nn = (long long)N * N;
    if (kStageOff * 8 + (long long)n_chains * nn * 8 > kScratchBytes)
        return free_error(DQMC_EINVAL, "dqmc_apply_bop_noalloc: n_chains * N * N * 8 bytes exceed the "
                                       "DQMC_NOALLOC_SCRATCH_BYTES scratch (less 8 KiB of constants)");
    double* S = nullptr;
    const int rc = scratch_for_current_device(&S);
    if (rc) return free_error(rc, "dqmc_apply_bop_noalloc: no device, or the scratch address is unavailable");
    hipStream_t st = (hipStream_t)stream;

    // the same host arithmetic as dqmc_create / dqmc_init, so the factors carry the same bits
    const double dt = p->beta / p->nt;
    const double alpha = std::acosh(std::exp(0.5 * dt * p->u));
    const double m = dt * p->mu;
    Block256 e, f, t;
    ring_expk16(nl, dt * p->nnt, e.v);
    ring_expk16(nl, -dt * p->nnt, f.v);
    t.v[0] = std::exp(m + alpha); t.v[1] = std::exp(m - alpha);
    t.v[2] = std::exp(-m - alpha); t.v[3] = std::exp(-m + alpha);
    hipLaunchKernelGGL(store_block_kernel, dim3(1), dim3(256), 0, st, e, S + kE, 256);
    hipLaunchKernelGGL(store_block_kernel, dim3(1), dim3(256), 0, st, f, S + kF, 256);
    hipLaunchKernelGGL(store_block_kernel, dim3(1), dim3(256), 0, st, t, S + kTab, 4);
    if (hipGetLastError() != hipSuccess) return free_error(DQMC_EHIP, "dqmc_apply_bop_noalloc: launch failed");

    // B(t) = diag(e^{dt mu + alpha s}) E(x)E: the operations of dqmc_apply_bop
    DiagArg tab, itab;
    tab.mode = DIAG_FIELD; tab.v = S + kTab; tab.vstride = 0; tab.f = d_fields; tab.fstride = N;
    itab = tab; itab.v = S + kITab;
    KronArgs<double> k;
    k.X = d_inout; k.sX = nn; k.C = S + kStageOff; k.sC = nn;
    k.nl = nl; k.N = N; k.batch = n_chains;
    hipError_t err;
    switch (op) {
        case 0: k.E = S + kE; k.rowd = tab; err = launch_kron_left<double>(k, st); break;     // B X
        case 1: k.E = S + kE; k.innd = tab; err = launch_kron_right<double>(k, st); break;    // X B
        case 2: k.E = S + kF; k.innd = itab; err = launch_kron_left<double>(k, st); break;    // B^{-1} X
        case 3: k.E = S + kF; k.cold = itab; err = launch_kron_right<double>(k, st); break;   // X B^{-1}
        default: k.E = S + kE; k.innd = tab; err = launch_kron_left<double>(k, st); break;    // B^T X
    }
    if (err == hipSuccess) err = launch_copy_d(S + kStageOff, nn, d_inout, nn, nn, n_chains, st);
    if (err != hipSuccess) return free_error(DQMC_EHIP, "dqmc_apply_bop_noalloc: launch failed");
    return DQMC_OK;
}

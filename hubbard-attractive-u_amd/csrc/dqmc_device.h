// dqmc_device.h -- device helpers shared by the gfx950 kernels of the DQMC sweep engine: the MFMA
// operand layouts, the chain-major block order, diagonal scalings and register-only cross-lane
// reductions.
#pragma once
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
//  Cross-lane reductions (DPP and the gfx950 permlane swaps).
// ---------------------------------------------------------------------------------------------

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

}  // namespace dqmc

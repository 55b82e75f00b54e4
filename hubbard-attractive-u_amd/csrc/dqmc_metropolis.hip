// dqmc_metropolis.hip -- Metropolis sweep of one slice with delayed rank-k Sherman-Morrison updates:
// one workgroup per chain, FP64 ratio from G_ii, minstd_rand0 uniforms by jump-ahead, the pending
// update flushed into G on MFMA inside the workgroup.
#include "dqmc_device.h"

namespace dqmc {

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

// The next wrap's left product by the workgroup of chain b after its slice (FP64, N = 256): each wave
// transforms 2 columns at a time as kron_left16_kernel does (the same per-element arithmetic), through
// its own 16 x 17 transpose buffer at the start of the LDS, which the pending update no longer uses.
template <int MT>
__device__ __forceinline__ void metro_left16(const MetroArgs<double>& a, int b, unsigned char* smem)
{
    double(*zt)[16][17] = reinterpret_cast<double(*)[16][17]>(smem);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int NWV = MT / 64, CG = 2;   // columns in flight per wave (the kernel's register budget)
    double ef[4], ef2[4];
    load_efrag<double>(a.lE, lane, ef, ef2);
    const double* G = a.G + (long long)b * a.sG;
    double* C = a.lout + (long long)b * a.sG;
    const int8_t* f = a.fields + ((long long)b * a.L + a.lt) * 256;
    const double dp = a.ltab[2 * b], dm = a.ltab[2 * b + 1];
    double ds[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) ds[s] = f[lane + 64 * s] > 0 ? dp : dm;
    const bool rows = a.lmode == 1;
    for (int c0 = wave * CG; c0 < 256; c0 += NWV * CG) {
        double vin[CG][4];
#pragma unroll
        for (int cc = 0; cc < CG; ++cc)
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const double v = G[(long long)(c0 + cc) * 256 + lane + 64 * s];
                vin[cc][s] = rows ? v : v * ds[s];
            }
#pragma unroll
        for (int cc = 0; cc < CG; ++cc) {
            d4_t z;
            kron16_vec<double>(ef, ef2, vin[cc], z);
#pragma unroll
            for (int r = 0; r < 4; ++r) zt[wave][lane & 15][Mfma<double>::row(lane, r)] = z[r];
            wave_lds_sync();
            double zv[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) zv[s] = zt[wave][(lane >> 4) + 4 * s][lane & 15];
            wave_lds_sync();   // read before the next column overwrites the buffer
#pragma unroll
            for (int s = 0; s < 4; ++s) C[(long long)(c0 + cc) * 256 + lane + 64 * s] = rows ? zv[s] * ds[s] : zv[s];
        }
    }
}

// The wrap's right product by the workgroup of chain b after its left product (FP64, N = 256): G =
// lout diag(s') (F(x)F) diag(c'), 64 rows of lout at a time staged row-major in LDS (row stride 273,
// one pad word per 16 sites, as kron_right16_kernel does for 32 rows), each row transformed by one
// wave; the same per-element arithmetic as kron_right16_kernel.
#define MR_ROWS 64
#define MR_LD 273
__device__ __forceinline__ int mr_pad(int site) { return site + (site >> 4); }
template <int MT>
__device__ __forceinline__ void metro_right16(const MetroArgs<double>& a, int b, unsigned char* smem)
{
    double* tile = reinterpret_cast<double*>(smem);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NWV = MT / 64;
    const double* X = a.lout + (long long)b * a.sG;
    double* C = a.G + (long long)b * a.sG;
    const int8_t* f = a.fields + ((long long)b * a.L + a.lt) * 256;
    const double dp = a.rtab[2 * b], dm = a.rtab[2 * b + 1];
    const bool inner = a.rmode == 2;
    double ef[4], ef2[4];
    load_efrag<double>(a.rE, lane, ef, ef2);
    for (int r0 = 0; r0 < 256; r0 += MR_ROWS) {
        for (int e = tid; e < MR_ROWS / 2 * 256; e += MT) {
            const int rr = 2 * (e & (MR_ROWS / 2 - 1)), c = e / (MR_ROWS / 2);
            const double2 v = *reinterpret_cast<const double2*>(X + r0 + rr + (long long)c * 256);
            double v0 = v.x, v1 = v.y;
            if (inner) {
                const double d = f[c] > 0 ? dp : dm;
                v0 = v0 * d; v1 = v1 * d;
            }
            tile[rr * MR_LD + mr_pad(c)] = v0;
            tile[(rr + 1) * MR_LD + mr_pad(c)] = v1;
        }
        __syncthreads();
        for (int q = wave; q < MR_ROWS; q += NWV) {
            double* row = tile + q * MR_LD;
            double vin[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) vin[s] = row[mr_pad(lane + 64 * s)];
            d4_t z;
            kron16_vec<double>(ef, ef2, vin, z);
#pragma unroll
            for (int r = 0; r < 4; ++r) row[mr_pad(Mfma<double>::row(lane, r) + 16 * (lane & 15))] = z[r];
        }
        __syncthreads();
        for (int e = tid; e < MR_ROWS / 2 * 256; e += MT) {
            const int rr = 2 * (e & (MR_ROWS / 2 - 1)), c = e / (MR_ROWS / 2);
            double v0 = tile[rr * MR_LD + mr_pad(c)], v1 = tile[(rr + 1) * MR_LD + mr_pad(c)];
            if (!inner) {
                const double d = f[c] > 0 ? dp : dm;
                v0 *= d; v1 *= d;
            }
            double2 v; v.x = v0; v.y = v1;
            *reinterpret_cast<double2*>(C + r0 + rr + (long long)c * 256) = v;
        }
        __syncthreads();   // the next rows overwrite the tile
    }
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
    const int b = a.chain0 + blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
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
    // wave 0 runs the latency-bound site loops beside the flushing waves: it issues first
    if (tid < 64) __builtin_amdgcn_s_setprio(3);
    if (tid < 64) metro_site_block(L, G, f, N, kb, 0, min(kb, N), dpos, dneg, health);
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6, nwv = nth >> 6, l16 = lane & 15;
    for (int b0 = 0; b0 < N; b0 += kb) {
        // The count of block b0 sits in its own parity slot: when na == 0 no barrier separates this
        // read from the site loop of the next block, which writes the other slot.
        const int na = L.nacc[(b0 / kb) & 1];
        nacc += na;
        const int b1 = b0 + kb, bs1 = (b1 < N) ? min(kb, N - b1) : 0;
        const bool next = b1 < N;
        const int tr0 = next ? b1 / 16 : 0, tr1 = next ? (b1 + bs1 - 1) / 16 : 0;
        const int a4 = ((na + 3) / 4) * 4;
        if (na > 0) {
            // Rebuild X (columns) and Y (rows) of the rank-na update from G0.  Each wave takes whole
            // 64-column tasks, X or Y, reads the G0 entries of its columns (a column segment of G for X,
            // one element of each of its 64 columns per row for Y) and runs the recursion of its columns
            // in registers with all 64 lanes active (the coefficients travel by DPP broadcast, which
            // needs every source lane).
            const int ys = L.ys, nct = (N + 63) / 64;
            const int qlane = (lane < na) ? L.ql[lane] : 0;   // lane m holds q_m
            const int qa = (l16 < na) ? L.ql[l16] : MB_MAX, qb = (16 + l16 < na) ? L.ql[16 + l16] : MB_MAX;
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
                        xr[m] = (m < na && jin) ? (double)G[(b0 + q) + (long long)j * N] - (j == b0 + q ? 1.0 : 0.0) : 0.0;
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
    if constexpr (sizeof(T) == 8) {
        // the final flush's stores of G are visible to the whole workgroup past the barrier above
        if (a.lmode != 0) {
            metro_left16<MT>(a, b, smem);
            if (a.rmode != 0) {
                __syncthreads();   // lout complete, and the left product's transpose buffers free
                metro_right16<MT>(a, b, smem);
            }
        }
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
    if (a.lmode != 0 && (sizeof(T) != 8 || a.N != 256 || !a.lout || !a.lE || !a.ltab || a.lt < 0 || a.lt >= a.L))
        return hipErrorInvalidValue;
    size_t shm = metropolis_lds_bytes(a.N, a.kmax, sizeof(T));
    if (a.lmode != 0 && shm < (size_t)16 * 16 * 17 * 8) shm = (size_t)16 * 16 * 17 * 8;   // the epilogue's transpose buffers
    if (a.rmode != 0) {
        if (a.lmode == 0 || !a.rE || !a.rtab) return hipErrorInvalidValue;
        if (shm < (size_t)MR_ROWS * MR_LD * 8) shm = (size_t)MR_ROWS * MR_LD * 8;   // the right product's row tile
    }
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

}  // namespace dqmc

/*
 * dqmc_oracle.c -- FP64 CPU restatement of the reference DQMC sweep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker and the CPU
 * baseline ("port") for the HIP engine.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Every routine follows the reference's operation order and cites the file:line
 * it restates (paths relative to the reference repository root):
 *
 *   RNG ............ src/random.cpp:14-15, libstdc++ minstd_rand0 +
 *                    generate_canonical<double,53> + bernoulli_distribution
 *   lattice / K .... src/square_lattice.cpp:30-61, src/hubbard.cpp:25-48
 *   B operations ... src/hubbard.cpp:148-156, 254-341; the is_fft products
 *                    src/hubbard.cpp:50-81, 159-246 (orc_set_fft)
 *   ratio / SM ..... src/hubbard.cpp:123-139
 *   SVD stack ...... include/utils/svd_stack.hpp:51-188
 *   numerics ....... include/utils/stable_numerics.hpp:38-324
 *   core sweeps .... src/dqmc_core.cpp:67-390
 *   observables .... include/observable_methods.hpp:39-195
 *
 * Third-party arithmetic the reference delegates (MKL dgesvd/dgemm, Eigen
 * PartialPivLU / FullPivHouseholderQR / MatrixFunctions::exp) is restated here
 * from the published algorithms:
 *   - SVD: one-sided (Hestenes) Jacobi SVD, singular values sorted descending,
 *     V returned (not V^T) as linear_algebra.hpp:37,86 does.
 *   - inverse: LU with partial (row) pivoting, then solve against I
 *     (Eigen 3.4 PartialPivLU::inverse).
 *   - full-pivot Householder QR solve (Eigen 3.4 FullPivHouseholderQR::solve):
 *     at step k the largest |a_ij| of the trailing block is moved to (k,k) by a
 *     row and a column transposition, then a Householder reflector zeroes the
 *     column below the diagonal.
 *   - exp(-dt K): K is real symmetric; K = W diag(l) W^T by cyclic Jacobi, then
 *     exp(-dt K) = W diag(exp(-dt l)) W^T.
 *   - is_fft (MKL DFTI, fft_solver.hpp:20-91): a direct 2-D DFT of length nl per
 *     axis, with the full complex spectrum.  The reference asks MKL for a real
 *     transform in the conjugate-even packed layout (fft_solver.hpp:60-61), which
 *     holds only nl x (nl/2+1) frequencies; the restatement keeps the operator
 *     the reference's comments state, (inv)expK = F^-1 diag(e^{-+dt E(k)}) F
 *     (src/hubbard.cpp:174-178).
 * When MKL's runtime (libmkl_rt) can be dlopen'ed, orc_set_mkl(1) routes the
 * GEMMs and the SVD through cblas_dgemm / LAPACKE_dgesvd exactly as the
 * reference does (EIGEN_USE_MKL_ALL, linear_algebra.hpp:74), and the
 * eigendecomposition behind exp(-dt K) through LAPACKE_dsyev; this times a
 * realistic CPU baseline and keeps the checks on large lattices fast.
 */
#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, n) ((size_t)(i) + (size_t)(j) * (size_t)(n))

/* ------------------------------------------------------------------------- */
/*  MKL (optional)                                                             */
/* ------------------------------------------------------------------------- */
typedef void (*cblas_dgemm_t)(int, int, int, int, int, int, double, const double*, int,
                              const double*, int, double, double*, int);
typedef int (*lapacke_dgesvd_t)(int, char, char, int, int, double*, int, double*, double*, int,
                                double*, int, double*);
typedef int (*lapacke_dsyev_t)(int, char, char, int, double*, int, double*);
typedef int (*mkl_set_layer_t)(int);
static cblas_dgemm_t p_dgemm = NULL;
static lapacke_dgesvd_t p_dgesvd = NULL;
static lapacke_dsyev_t p_dsyev = NULL;
static int g_use_mkl = 0;

int orc_set_mkl(int on)
{
    if (!on) { g_use_mkl = 0; return 0; }
    if (!p_dgemm) {
        void* h = dlopen("libmkl_rt.so", RTLD_NOW | RTLD_GLOBAL);
        if (!h) h = dlopen("/opt/conda/lib/libmkl_rt.so", RTLD_NOW | RTLD_GLOBAL);
        if (!h) return -1;
        mkl_set_layer_t setl = (mkl_set_layer_t)dlsym(h, "MKL_Set_Threading_Layer");
        if (setl) setl(1 /* MKL_THREADING_SEQUENTIAL */);
        p_dgemm = (cblas_dgemm_t)dlsym(h, "cblas_dgemm");
        p_dgesvd = (lapacke_dgesvd_t)dlsym(h, "LAPACKE_dgesvd");
        p_dsyev = (lapacke_dsyev_t)dlsym(h, "LAPACKE_dsyev");
        if (!p_dgemm || !p_dgesvd) { p_dgemm = NULL; p_dgesvd = NULL; return -2; }
    }
    g_use_mkl = 1;
    return 0;
}

/* ------------------------------------------------------------------------- */
/*  RNG: std::default_random_engine (= minstd_rand0 in libstdc++)             */
/*  src/random.cpp:14-15; libstdc++ bits/random.tcc:116-124 (seed),           */
/*  bits/random.tcc generate_canonical, random.h bernoulli_distribution.      */
/* ------------------------------------------------------------------------- */
typedef struct { uint64_t x; } orc_rng;
/* 1 (default): (x2-1)*R + (x1-1) as one fused multiply-add, which is what g++ -O3
 * -march=native (the reference build, CMakeLists.txt:5-6) emits on an FMA-capable host. */
static int g_canonical_fma = 1;

void orc_rng_set_fma(int on) { g_canonical_fma = on; }

static void rng_seed(orc_rng* r, uint64_t s)
{
    uint64_t v = s % 2147483647ULL;
    r->x = v ? v : 1ULL;
}
static uint64_t rng_next(orc_rng* r)
{
    r->x = (r->x * 16807ULL) % 2147483647ULL;
    return r->x;
}
/* generate_canonical<double,53>(minstd_rand0): two draws, R = max-min+1. */
static double rng_canonical(orc_rng* r)
{
    const double R = 2147483646.0;
    const double R2 = (double)(2147483646.0L * 2147483646.0L);
    double sum = (double)(rng_next(r) - 1ULL);
    double x2 = (double)(rng_next(r) - 1ULL);
    if (g_canonical_fma) sum = fma(x2, R, sum);
    else { volatile double prod = x2 * R; sum = sum + prod; }
    double ret = sum / R2;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}
static int rng_bernoulli(orc_rng* r, double p) { return rng_canonical(r) < p; }

/* exported RNG helpers for the known-answer tests */
void orc_rng_stream(uint64_t seed, int n, uint64_t* out)
{
    orc_rng r; rng_seed(&r, seed);
    for (int i = 0; i < n; ++i) out[i] = rng_next(&r);
}
void orc_rng_canonical_stream(uint64_t seed, int n, double* out)
{
    orc_rng r; rng_seed(&r, seed);
    for (int i = 0; i < n; ++i) out[i] = rng_canonical(&r);
}

/* ------------------------------------------------------------------------- */
/*  Dense helpers (column-major, n x n)                                        */
/* ------------------------------------------------------------------------- */
static double* mat_alloc(int n) { return (double*)calloc((size_t)n * n, sizeof(double)); }
static void mat_eye(int n, double* a)
{
    memset(a, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) a[IDX(i, i, n)] = 1.0;
}
static void mat_copy(int n, double* dst, const double* src) { memcpy(dst, src, sizeof(double) * (size_t)n * n); }
static void mat_transpose(int n, double* dst, const double* src)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) dst[IDX(j, i, n)] = src[IDX(i, j, n)];
}

/* C = op(A) * op(B); ta/tb: 0 = no transpose, 1 = transpose. C must not alias. */
static void gemm(int n, int ta, const double* A, int tb, const double* B, double* C)
{
    if (g_use_mkl && p_dgemm) {
        p_dgemm(102 /*ColMajor*/, ta ? 112 : 111, tb ? 112 : 111, n, n, n, 1.0, A, n, B, n, 0.0, C, n);
        return;
    }
    memset(C, 0, sizeof(double) * (size_t)n * n);
    double* At = NULL;
    const double* Aa = A;
    if (ta) { At = mat_alloc(n); mat_transpose(n, At, A); Aa = At; }
    for (int j = 0; j < n; ++j) {
        double* c = C + (size_t)j * n;
        for (int k = 0; k < n; ++k) {
            const double b = tb ? B[IDX(j, k, n)] : B[IDX(k, j, n)];
            const double* a = Aa + (size_t)k * n;
            for (int i = 0; i < n; ++i) c[i] += a[i] * b;
        }
    }
    free(At);
}

/* ------------------------------------------------------------------------- */
/*  Symmetric eigen-decomposition (cyclic Jacobi) and exp(-dt K)               */
/*  restates Eigen MatrixFunctions exp() used at src/hubbard.cpp:46-47        */
/* ------------------------------------------------------------------------- */
static void sym_eig_jacobi(int n, double* a /* destroyed */, double* w, double* v)
{
    mat_eye(n, v);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double x = a[IDX(i, j, n)] * a[IDX(i, j, n)];
                tot += x;
                if (i != j) off += x;
            }
        if (off <= 1e-34 * (tot > 0 ? tot : 1.0)) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = a[IDX(p, q, n)];
                if (fabs(apq) < 1e-300) continue;
                const double app = a[IDX(p, p, n)], aqq = a[IDX(q, q, n)];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) { /* columns p,q */
                    const double akp = a[IDX(k, p, n)], akq = a[IDX(k, q, n)];
                    a[IDX(k, p, n)] = c * akp - s * akq;
                    a[IDX(k, q, n)] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) { /* rows p,q */
                    const double apk = a[IDX(p, k, n)], aqk = a[IDX(q, k, n)];
                    a[IDX(p, k, n)] = c * apk - s * aqk;
                    a[IDX(q, k, n)] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = v[IDX(k, p, n)], vkq = v[IDX(k, q, n)];
                    v[IDX(k, p, n)] = c * vkp - s * vkq;
                    v[IDX(k, q, n)] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = a[IDX(i, i, n)];
}

/* K = v diag(w) v^T for symmetric K: LAPACKE_dsyev when MKL is loaded (large lattices), else the
 * cyclic Jacobi iteration above */
static void sym_eig(int n, const double* K, double* w, double* v)
{
    if (g_use_mkl && p_dsyev) {
        mat_copy(n, v, K);
        if (p_dsyev(102 /*ColMajor*/, 'V', 'U', n, v, n, w) == 0) return;
    }
    double* a = mat_alloc(n);
    mat_copy(n, a, K);
    sym_eig_jacobi(n, a, w, v);
    free(a);
}

/* out = v diag(exp(scale w)) v^T, each entry summed over k in order */
static void sym_expm_from(int n, const double* w, const double* v, double scale, double* out)
{
    double* e = (double*)malloc(sizeof(double) * n);
    double* vr = mat_alloc(n);   /* row-major copy: vr[i n + k] = v(i, k) */
    for (int k = 0; k < n; ++k) e[k] = exp(scale * w[k]);
    mat_transpose(n, vr, v);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const double* vi = vr + (size_t)i * n;
            const double* vj = vr + (size_t)j * n;
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += vi[k] * e[k] * vj[k];
            out[IDX(i, j, n)] = s;
        }
    free(e); free(vr);
}

/* exported for tests: nearest-neighbour adjacency of the periodic nl x nl lattice
 * (src/square_lattice.cpp:30-61; entries accumulate, so nl=2 gives 2 and nl=1 gives 4). */
void orc_nn_hoppings(int nl, double* A)
{
    const int n = nl * nl;
    memset(A, 0, sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) {
        const int x = i % nl, y = i / nl;
        const int xp = ((x + 1) % nl) + nl * y;
        const int yp = x + nl * ((y + 1) % nl);
        A[IDX(i, xp, n)] += 1.0; A[IDX(xp, i, n)] += 1.0;
        A[IDX(i, yp, n)] += 1.0; A[IDX(yp, i, n)] += 1.0;
    }
}

/* exported: expK = exp(-dt K), invExpK = exp(+dt K), K = -nnt*A - mu*I (src/hubbard.cpp:41-47) */
void orc_expk(int nl, double dt, double nnt, double mu, double* expK, double* invExpK)
{
    const int n = nl * nl;
    double* K = mat_alloc(n);
    orc_nn_hoppings(nl, K);
    for (size_t e = 0; e < (size_t)n * n; ++e) K[e] *= -nnt;
    for (int i = 0; i < n; ++i) K[IDX(i, i, n)] -= mu;
    double* v = mat_alloc(n);
    double* w = (double*)malloc(sizeof(double) * n);
    sym_eig(n, K, w, v);
    if (expK) sym_expm_from(n, w, v, -dt, expK);
    if (invExpK) sym_expm_from(n, w, v, +dt, invExpK);
    free(K); free(v); free(w);
}

/* ------------------------------------------------------------------------- */
/*  SVD (one-sided Jacobi), restating LinearAlgebra::mkl_lapack_dgesvd         */
/*  (include/utils/linear_algebra.hpp:46-99): A = U diag(s) V^T, s descending, */
/*  V returned (not V^T).                                                      */
/* ------------------------------------------------------------------------- */
static void svd_jacobi(int n, const double* A, double* U, double* s, double* V)
{
    double* W = mat_alloc(n);
    double* Vt = mat_alloc(n);
    mat_copy(n, W, A);
    mat_eye(n, Vt);
    const double tol = 1e-15;
    for (int sweep = 0; sweep < 80; ++sweep) {
        int rotated = 0;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double al = 0, be = 0, ga = 0;
                const double* wp = W + (size_t)p * n;
                const double* wq = W + (size_t)q * n;
                for (int k = 0; k < n; ++k) { al += wp[k] * wp[k]; be += wq[k] * wq[k]; ga += wp[k] * wq[k]; }
                if (ga == 0.0 || fabs(ga) <= tol * sqrt(al * be)) continue;
                rotated = 1;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
                double* wpp = W + (size_t)p * n;
                double* wqq = W + (size_t)q * n;
                double* vp = Vt + (size_t)p * n;
                double* vq = Vt + (size_t)q * n;
                for (int k = 0; k < n; ++k) {
                    const double x = wpp[k], y = wqq[k];
                    wpp[k] = c * x - sn * y; wqq[k] = sn * x + c * y;
                    const double vx = vp[k], vy = vq[k];
                    vp[k] = c * vx - sn * vy; vq[k] = sn * vx + c * vy;
                }
            }
        if (!rotated) break;
    }
    /* singular values = column norms; sort descending */
    int* ord = (int*)malloc(sizeof(int) * n);
    double* nrm = (double*)malloc(sizeof(double) * n);
    for (int j = 0; j < n; ++j) {
        double t = 0;
        for (int k = 0; k < n; ++k) t += W[IDX(k, j, n)] * W[IDX(k, j, n)];
        nrm[j] = sqrt(t); ord[j] = j;
    }
    for (int a = 1; a < n; ++a) { /* insertion sort, stable */
        int key = ord[a]; int b = a - 1;
        while (b >= 0 && nrm[ord[b]] < nrm[key]) { ord[b + 1] = ord[b]; --b; }
        ord[b + 1] = key;
    }
    for (int j = 0; j < n; ++j) {
        const int o = ord[j];
        s[j] = nrm[o];
        for (int k = 0; k < n; ++k) {
            U[IDX(k, j, n)] = nrm[o] > 0 ? W[IDX(k, o, n)] / nrm[o] : (k == j ? 1.0 : 0.0);
            V[IDX(k, j, n)] = Vt[IDX(k, o, n)];
        }
    }
    free(ord); free(nrm); free(W); free(Vt);
}

static int svd(int n, const double* A, double* U, double* s, double* V)
{
    if (g_use_mkl && p_dgesvd) {
        double* a = mat_alloc(n);
        double* vt = mat_alloc(n);
        double* superb = (double*)malloc(sizeof(double) * (n > 1 ? n - 1 : 1));
        mat_copy(n, a, A);
        int info = p_dgesvd(102, 'S', 'S', n, n, a, n, s, U, n, vt, n, superb);
        mat_transpose(n, V, vt);
        free(a); free(vt); free(superb);
        return info > 0 ? -1 : 0;
    }
    svd_jacobi(n, A, U, s, V);
    return 0;
}

/* exported for tests */
int orc_svd(int n, const double* A, double* U, double* s, double* V) { return svd(n, A, U, s, V); }

/* ------------------------------------------------------------------------- */
/*  LU with partial pivoting + inverse (Eigen PartialPivLU::inverse)           */
/*  used at include/utils/stable_numerics.hpp:235, 304, 322                    */
/* ------------------------------------------------------------------------- */
static int lu_inverse(int n, const double* A, double* Ainv)
{
    double* lu = mat_alloc(n);
    int* piv = (int*)malloc(sizeof(int) * n);
    mat_copy(n, lu, A);
    int sing = 0;
    for (int k = 0; k < n; ++k) {
        int p = k; double mx = fabs(lu[IDX(k, k, n)]);
        for (int i = k + 1; i < n; ++i) if (fabs(lu[IDX(i, k, n)]) > mx) { mx = fabs(lu[IDX(i, k, n)]); p = i; }
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) { double t = lu[IDX(k, j, n)]; lu[IDX(k, j, n)] = lu[IDX(p, j, n)]; lu[IDX(p, j, n)] = t; }
        const double d = lu[IDX(k, k, n)];
        if (d == 0.0) { sing = 1; continue; }
        for (int i = k + 1; i < n; ++i) lu[IDX(i, k, n)] /= d;
        for (int j = k + 1; j < n; ++j) {
            const double f = lu[IDX(k, j, n)];
            if (f == 0.0) continue;
            for (int i = k + 1; i < n; ++i) lu[IDX(i, j, n)] -= lu[IDX(i, k, n)] * f;
        }
    }
    /* solve LU X = P I */
    mat_eye(n, Ainv);
    for (int k = 0; k < n; ++k)
        if (piv[k] != k)
            for (int j = 0; j < n; ++j) { double t = Ainv[IDX(k, j, n)]; Ainv[IDX(k, j, n)] = Ainv[IDX(piv[k], j, n)]; Ainv[IDX(piv[k], j, n)] = t; }
    for (int j = 0; j < n; ++j) {
        double* x = Ainv + (size_t)j * n;
        for (int i = 0; i < n; ++i) { double s = x[i]; for (int k = 0; k < i; ++k) s -= lu[IDX(i, k, n)] * x[k]; x[i] = s; }
        for (int i = n - 1; i >= 0; --i) { double s = x[i]; for (int k = i + 1; k < n; ++k) s -= lu[IDX(i, k, n)] * x[k]; x[i] = s / lu[IDX(i, i, n)]; }
    }
    free(lu); free(piv);
    return sing ? -1 : 0;
}
int orc_inverse(int n, const double* A, double* Ainv) { return lu_inverse(n, A, Ainv); }

/* ------------------------------------------------------------------------- */
/*  Full-pivot Householder QR solve X = H^{-1} B (Eigen FullPivHouseholderQR)  */
/*  used at include/utils/stable_numerics.hpp:154                              */
/* ------------------------------------------------------------------------- */
static void fullpiv_qr_solve(int n, const double* H, const double* B, double* X)
{
    double* a = mat_alloc(n);
    double* b = mat_alloc(n);
    int* rowt = (int*)malloc(sizeof(int) * n);
    int* colp = (int*)malloc(sizeof(int) * n);
    double* hc = (double*)malloc(sizeof(double) * n);
    mat_copy(n, a, H);
    mat_copy(n, b, B);
    for (int j = 0; j < n; ++j) colp[j] = j;
    for (int k = 0; k < n; ++k) {
        int pi = k, pj = k; double mx = -1.0;
        for (int j = k; j < n; ++j)
            for (int i = k; i < n; ++i) {
                const double v = fabs(a[IDX(i, j, n)]);
                if (v > mx) { mx = v; pi = i; pj = j; }
            }
        rowt[k] = pi;
        if (pi != k) {
            for (int j = 0; j < n; ++j) { double t = a[IDX(k, j, n)]; a[IDX(k, j, n)] = a[IDX(pi, j, n)]; a[IDX(pi, j, n)] = t; }
            for (int j = 0; j < n; ++j) { double t = b[IDX(k, j, n)]; b[IDX(k, j, n)] = b[IDX(pi, j, n)]; b[IDX(pi, j, n)] = t; }
        }
        if (pj != k) {
            for (int i = 0; i < n; ++i) { double t = a[IDX(i, k, n)]; a[IDX(i, k, n)] = a[IDX(i, pj, n)]; a[IDX(i, pj, n)] = t; }
            int t = colp[k]; colp[k] = colp[pj]; colp[pj] = t;
        }
        /* Householder on column k, rows k..n-1 (LAPACK dlarfg convention) */
        double alpha = a[IDX(k, k, n)], sig = 0.0;
        for (int i = k + 1; i < n; ++i) sig += a[IDX(i, k, n)] * a[IDX(i, k, n)];
        if (sig == 0.0) { hc[k] = 0.0; continue; }
        const double beta = -copysign(sqrt(alpha * alpha + sig), alpha);
        const double tau = (beta - alpha) / beta;
        const double scal = 1.0 / (alpha - beta);
        for (int i = k + 1; i < n; ++i) a[IDX(i, k, n)] *= scal;
        a[IDX(k, k, n)] = beta;
        hc[k] = tau;
        /* apply H = I - tau v v^T to trailing columns of a and to b */
        for (int j = k + 1; j < n; ++j) {
            double w = a[IDX(k, j, n)];
            for (int i = k + 1; i < n; ++i) w += a[IDX(i, k, n)] * a[IDX(i, j, n)];
            w *= tau;
            a[IDX(k, j, n)] -= w;
            for (int i = k + 1; i < n; ++i) a[IDX(i, j, n)] -= w * a[IDX(i, k, n)];
        }
        for (int j = 0; j < n; ++j) {
            double w = b[IDX(k, j, n)];
            for (int i = k + 1; i < n; ++i) w += a[IDX(i, k, n)] * b[IDX(i, j, n)];
            w *= tau;
            b[IDX(k, j, n)] -= w;
            for (int i = k + 1; i < n; ++i) b[IDX(i, j, n)] -= w * a[IDX(i, k, n)];
        }
    }
    /* back substitution R y = Q^T P b, then X = Pc y */
    for (int j = 0; j < n; ++j) {
        double* y = b + (size_t)j * n;
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < n; ++k) s -= a[IDX(i, k, n)] * y[k];
            y[i] = s / a[IDX(i, i, n)];
        }
        for (int i = 0; i < n; ++i) X[IDX(colp[i], j, n)] = y[i];
    }
    free(a); free(b); free(rowt); free(colp); free(hc);
}
void orc_fullpiv_qr_solve(int n, const double* H, const double* B, double* X) { fullpiv_qr_solve(n, H, B, X); }

/* ------------------------------------------------------------------------- */
/*  SVD stack (include/utils/svd_stack.hpp:51-188)                             */
/* ------------------------------------------------------------------------- */
typedef struct {
    int n, cap, len;
    double** U; double** V; double** S;
    double* tmp;
} orc_stack;

static void stack_init(orc_stack* st, int n, int cap)
{
    st->n = n; st->cap = cap; st->len = 0;
    st->U = (double**)malloc(sizeof(double*) * cap);
    st->V = (double**)malloc(sizeof(double*) * cap);
    st->S = (double**)malloc(sizeof(double*) * cap);
    for (int i = 0; i < cap; ++i) { st->U[i] = mat_alloc(n); st->V[i] = mat_alloc(n); st->S[i] = (double*)calloc(n, sizeof(double)); }
    st->tmp = mat_alloc(n);
}
static void stack_free(orc_stack* st)
{
    for (int i = 0; i < st->cap; ++i) { free(st->U[i]); free(st->V[i]); free(st->S[i]); }
    free(st->U); free(st->V); free(st->S); free(st->tmp);
}
/* push: svd_stack.hpp:107-175 ; tmp = (A * U_top) * diag(S_top) (:144) */
static int stack_push(orc_stack* st, const double* A)
{
    const int n = st->n;
    if (st->len >= st->cap) return -1;
    int rc;
    if (st->len == 0) {
        rc = svd(n, A, st->U[0], st->S[0], st->V[0]);
    } else {
        const int top = st->len - 1;
        gemm(n, 0, A, 0, st->U[top], st->tmp);
        for (int j = 0; j < n; ++j) {
            const double s = st->S[top][j];
            for (int i = 0; i < n; ++i) st->tmp[IDX(i, j, n)] *= s;
        }
        rc = svd(n, st->tmp, st->U[st->len], st->S[st->len], st->V[st->len]);
    }
    st->len += 1;
    return rc;
}
static void stack_pop(orc_stack* st) { if (st->len > 0) st->len -= 1; }
/* MatrixV(): V_0 * V_1 * ... * V_{len-1} (svd_stack.hpp:95-103) */
static void stack_matV(orc_stack* st, double* out)
{
    const int n = st->n;
    double* t = mat_alloc(n);
    mat_copy(n, out, st->V[0]);
    for (int i = 1; i < st->len; ++i) { gemm(n, 0, out, 0, st->V[i], t); mat_copy(n, out, t); }
    free(t);
}

/* ------------------------------------------------------------------------- */
/*  StableNumerics (include/utils/stable_numerics.hpp)                         */
/* ------------------------------------------------------------------------- */
/* matrix_compare_error (:38-48): signed max of (mat1 - mat2) */
static double compare_error(int n, const double* a, const double* b)
{
    double e = -INFINITY;
    for (size_t k = 0; k < (size_t)n * n; ++k) { const double d = a[k] - b[k]; if (d > e) e = d; }
    return e;
}
/* div_dvec_max_min (:56-64) */
static void div_max_min(int n, const double* d, double* dmax, double* dmin)
{
    for (int i = 0; i < n; ++i) { dmax[i] = d[i] > 1.0 ? d[i] : 1.0; dmin[i] = d[i] < 1.0 ? d[i] : 1.0; }
}
/* mult_v_invd_u (:71-97): z(j,i) = sum_k v(j,k)*u(k,i)/d(k); naive loops as in the reference */
static void mult_v_invd_u(int n, const double* v, const double* d, const double* u, double* z)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double t = 0.0;
            for (int k = 0; k < n; ++k) t += v[IDX(j, k, n)] * u[IDX(k, i, n)] / d[k];
            z[IDX(j, i, n)] = t;
        }
}
/* mult_v_d_u (:104-130) */
static void mult_v_d_u(int n, const double* v, const double* d, const double* u, double* z)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double t = 0.0;
            for (int k = 0; k < n; ++k) t += v[IDX(j, k, n)] * u[IDX(k, i, n)] * d[k];
            z[IDX(j, i, n)] = t;
        }
}
/* compute_gf_00bb (:136-155): (1 + U S V^T)^{-1} via full-pivot QR */
static void gf_00bb(int n, const double* u, const double* s, const double* v, double* g)
{
    double* sl = (double*)malloc(sizeof(double) * n);
    double* sr = (double*)malloc(sizeof(double) * n);
    double* H = mat_alloc(n);
    double* R = mat_alloc(n);
    for (int i = 0; i < n; ++i) { if (s[i] > 1.0) { sl[i] = 1.0 / s[i]; sr[i] = 1.0; } else { sl[i] = 1.0; sr[i] = s[i]; } }
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            R[IDX(i, j, n)] = sl[i] * u[IDX(j, i, n)];
            H[IDX(i, j, n)] = sl[i] * u[IDX(j, i, n)] + sr[i] * v[IDX(j, i, n)];
        }
    fullpiv_qr_solve(n, H, R, g);
    free(sl); free(sr); free(H); free(R);
}

typedef struct {
    double *ul, *vl, *ur, *vr, *dl, *dr, *dlmax, *dlmin, *drmax, *drmin, *tA, *tB, *tC, *tD, *tmp, *inv, *neg;
} orc_work;

static void work_init(orc_work* w, int n)
{
    w->ul = mat_alloc(n); w->vl = mat_alloc(n); w->ur = mat_alloc(n); w->vr = mat_alloc(n);
    w->dl = calloc(n, sizeof(double)); w->dr = calloc(n, sizeof(double));
    w->dlmax = calloc(n, sizeof(double)); w->dlmin = calloc(n, sizeof(double));
    w->drmax = calloc(n, sizeof(double)); w->drmin = calloc(n, sizeof(double));
    w->tA = mat_alloc(n); w->tB = mat_alloc(n); w->tC = mat_alloc(n); w->tD = mat_alloc(n);
    w->tmp = mat_alloc(n); w->inv = mat_alloc(n); w->neg = mat_alloc(n);
}
static void work_free(orc_work* w)
{
    free(w->ul); free(w->vl); free(w->ur); free(w->vr); free(w->dl); free(w->dr);
    free(w->dlmax); free(w->dlmin); free(w->drmax); free(w->drmin);
    free(w->tA); free(w->tB); free(w->tC); free(w->tD); free(w->tmp); free(w->inv); free(w->neg);
}

static void load_lr(orc_work* w, int n, orc_stack* L, orc_stack* R)
{
    mat_copy(n, w->ul, L->U[L->len - 1]); stack_matV(L, w->vl); memcpy(w->dl, L->S[L->len - 1], sizeof(double) * n);
    mat_copy(n, w->ur, R->U[R->len - 1]); stack_matV(R, w->vr); memcpy(w->dr, R->S[R->len - 1], sizeof(double) * n);
    div_max_min(n, w->dl, w->dlmax, w->dlmin);
    div_max_min(n, w->dr, w->drmax, w->drmin);
}

/* compute_equaltime_gf (:186-239) */
static int equaltime_gf(orc_work* w, int n, orc_stack* L, orc_stack* R, double* gtt)
{
    if (L->len == 0) {
        stack_matV(R, w->vr);
        gf_00bb(n, w->vr, R->S[R->len - 1], R->U[R->len - 1], gtt);
        return 0;
    }
    if (R->len == 0) {
        stack_matV(L, w->vl);
        gf_00bb(n, L->U[L->len - 1], L->S[L->len - 1], w->vl, gtt);
        return 0;
    }
    load_lr(w, n, L, R);
    gemm(n, 1, w->ul, 0, w->ur, w->tA);
    gemm(n, 1, w->vl, 0, w->vr, w->tB);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            w->tA[IDX(i, j, n)] = w->tA[IDX(i, j, n)] / (w->dlmax[i] * w->drmax[j]);
            w->tB[IDX(i, j, n)] = w->tB[IDX(i, j, n)] * w->dlmin[i] * w->drmin[j];
        }
    for (size_t k = 0; k < (size_t)n * n; ++k) w->tmp[k] = w->tA[k] + w->tB[k];
    int rc = lu_inverse(n, w->tmp, w->inv);
    mult_v_invd_u(n, w->ur, w->drmax, w->inv, w->tA);
    mat_transpose(n, w->tC, w->ul);
    mult_v_invd_u(n, w->tA, w->dlmax, w->tC, gtt);
    return rc;
}

/* compute_dynamic_gf (:245-324) */
static int dynamic_gf(orc_work* w, int n, orc_stack* L, orc_stack* R, double* gt0, double* g0t, double* gtt)
{
    if (L->len == 0) {
        stack_matV(R, w->vr);
        gf_00bb(n, w->vr, R->S[R->len - 1], R->U[R->len - 1], gtt);
        mat_copy(n, gt0, gtt);
        for (size_t k = 0; k < (size_t)n * n; ++k) g0t[k] = gtt[k];
        for (int i = 0; i < n; ++i) g0t[IDX(i, i, n)] -= 1.0;
        return 0;
    }
    if (R->len == 0) {
        stack_matV(L, w->vl);
        gf_00bb(n, L->U[L->len - 1], L->S[L->len - 1], w->vl, gtt);
        for (size_t k = 0; k < (size_t)n * n; ++k) { gt0[k] = -gtt[k]; g0t[k] = -gtt[k]; }
        for (int i = 0; i < n; ++i) gt0[IDX(i, i, n)] += 1.0;
        return 0;
    }
    load_lr(w, n, L, R);
    gemm(n, 1, w->ul, 0, w->ur, w->tA);
    gemm(n, 1, w->vl, 0, w->vr, w->tB);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            w->tA[IDX(i, j, n)] = w->tA[IDX(i, j, n)] / (w->dlmax[i] * w->drmax[j]);
            w->tB[IDX(i, j, n)] = w->tB[IDX(i, j, n)] * w->dlmin[i] * w->drmin[j];
        }
    for (size_t k = 0; k < (size_t)n * n; ++k) w->tmp[k] = w->tA[k] + w->tB[k];
    int rc = lu_inverse(n, w->tmp, w->inv);
    mult_v_invd_u(n, w->ur, w->drmax, w->inv, w->tA);
    mat_transpose(n, w->tC, w->ul);
    mult_v_invd_u(n, w->tA, w->dlmax, w->tC, gtt);
    mat_transpose(n, w->tC, w->vl);
    mult_v_d_u(n, w->tA, w->dlmin, w->tC, gt0);

    gemm(n, 1, w->vr, 0, w->vl, w->tC);
    gemm(n, 1, w->ur, 0, w->ul, w->tD);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            w->tC[IDX(i, j, n)] = w->tC[IDX(i, j, n)] / (w->drmax[i] * w->dlmax[j]);
            w->tD[IDX(i, j, n)] = w->tD[IDX(i, j, n)] * w->drmin[i] * w->dlmin[j];
        }
    for (size_t k = 0; k < (size_t)n * n; ++k) w->tmp[k] = w->tC[k] + w->tD[k];
    rc |= lu_inverse(n, w->tmp, w->inv);
    for (size_t k = 0; k < (size_t)n * n; ++k) w->neg[k] = -w->vl[k];
    mult_v_invd_u(n, w->neg, w->dlmax, w->inv, w->tC);
    mat_transpose(n, w->tD, w->ur);
    mult_v_d_u(n, w->tC, w->drmin, w->tD, g0t);
    return rc;
}

/* ------------------------------------------------------------------------- */
/*  Model + Core                                                               */
/* ------------------------------------------------------------------------- */
typedef struct {
    int nl, n, nt, pace;
    double dt, beta, nnt, u, mu, alpha;
    int eqtime, dynamic, thermalization;
    int* fields;            /* [t][i] */
    double *expK, *invExpK;
    orc_rng rng;
    orc_stack left, right;
    orc_work w;
    double *gftt, *gft0, *gf0t, *stacks, *t1, *t2, *t3, *t4, *col, *row;
    double *vecgftt, *vecgft0, *vecgf0t; /* [t][n*n] */
    int current_t;
    double wrap_eq, wrap_dyn;
    long long accepts, visits;
    uint32_t* accmask;  /* [t][(n+31)/32]: accept bits of the last updating sweep */
    /* is_fft products: eigenvalues e^{-+dt E(k)} of (inv)expK at k = kx + nl*ky, twiddles
     * cos/sin(2 pi m / nl) at m, and complex work columns */
    int is_fft;
    double *lam_f, *lam_i, *tw_c, *tw_s, *wre, *wim, *vre, *vim;
} orc_ctx;

orc_ctx* orc_create(int nl, int nt, double beta, double nnt, double u, double mu, int pace, int eqtime, int dynamic)
{
    orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
    c->nl = nl; c->n = nl * nl; c->nt = nt; c->pace = pace; c->beta = beta; c->dt = beta / nt;
    c->nnt = nnt; c->u = u; c->mu = mu;
    c->alpha = acosh(exp(0.5 * c->dt * u));        /* src/hubbard.cpp:35 */
    c->eqtime = eqtime; c->dynamic = dynamic;
    const int n = c->n;
    c->fields = (int*)calloc((size_t)nt * n, sizeof(int));
    c->expK = mat_alloc(n); c->invExpK = mat_alloc(n);
    orc_expk(nl, c->dt, nnt, mu, c->expK, c->invExpK);
    rng_seed(&c->rng, 0);
    stack_init(&c->left, n, nt);                   /* dqmc_core.cpp:39-40: capacity nt */
    stack_init(&c->right, n, nt);
    work_init(&c->w, n);
    c->gftt = mat_alloc(n); c->gft0 = mat_alloc(n); c->gf0t = mat_alloc(n);
    c->stacks = mat_alloc(n); c->t1 = mat_alloc(n); c->t2 = mat_alloc(n); c->t3 = mat_alloc(n); c->t4 = mat_alloc(n);
    c->col = (double*)calloc(n, sizeof(double)); c->row = (double*)calloc(n, sizeof(double));
    if (eqtime || dynamic) c->vecgftt = (double*)calloc((size_t)nt * n * n, sizeof(double));
    if (dynamic) {
        c->vecgft0 = (double*)calloc((size_t)nt * n * n, sizeof(double));
        c->vecgf0t = (double*)calloc((size_t)nt * n * n, sizeof(double));
    }
    c->accmask = (uint32_t*)calloc((size_t)nt * ((n + 31) / 32), sizeof(uint32_t));
    return c;
}

void orc_destroy(orc_ctx* c)
{
    if (!c) return;
    free(c->fields); free(c->expK); free(c->invExpK);
    stack_free(&c->left); stack_free(&c->right); work_free(&c->w);
    free(c->gftt); free(c->gft0); free(c->gf0t); free(c->stacks); free(c->t1); free(c->t2); free(c->t3); free(c->t4);
    free(c->col); free(c->row);
    free(c->vecgftt); free(c->vecgft0); free(c->vecgf0t); free(c->accmask);
    free(c->lam_f); free(c->lam_i); free(c->tw_c); free(c->tw_s); free(c->wre); free(c->wim); free(c->vre); free(c->vim);
    free(c);
}

void orc_set_seed(orc_ctx* c, uint64_t seed) { rng_seed(&c->rng, seed); }
uint64_t orc_get_rng_state(orc_ctx* c) { return c->rng.x; }
void orc_set_rng_state(orc_ctx* c, uint64_t x) { c->rng.x = x; }
void orc_set_thermalization(orc_ctx* c, int th) { c->thermalization = th; }
double orc_alpha(orc_ctx* c) { return c->alpha; }

/* set_ising_fields_to_random (src/hubbard.cpp:102-110): t-major, bernoulli(0.5) -> +1/-1 */
void orc_fields_random(orc_ctx* c)
{
    for (int t = 0; t < c->nt; ++t)
        for (int i = 0; i < c->n; ++i) c->fields[t * c->n + i] = rng_bernoulli(&c->rng, 0.5) ? +1 : -1;
}
void orc_set_fields(orc_ctx* c, const int* s) { memcpy(c->fields, s, sizeof(int) * (size_t)c->nt * c->n); }
void orc_get_fields(orc_ctx* c, int* s) { memcpy(s, c->fields, sizeof(int) * (size_t)c->nt * c->n); }
void orc_get_expk(orc_ctx* c, double* e, double* ie) { if (e) mat_copy(c->n, e, c->expK); if (ie) mat_copy(c->n, ie, c->invExpK); }

/* expV scalings (src/hubbard.cpp:254-285) */
static void scale_rows(orc_ctx* c, double* g, int t, double sgn)
{
    const int n = c->n;
    for (int i = 0; i < n; ++i) {
        const double f = exp(sgn * c->alpha * c->fields[t * n + i]);
        for (int j = 0; j < n; ++j) g[IDX(i, j, n)] *= f;
    }
}
static void scale_cols(orc_ctx* c, double* g, int t, double sgn)
{
    const int n = c->n;
    for (int j = 0; j < n; ++j) {
        const double f = exp(sgn * c->alpha * c->fields[t * n + j]);
        for (int i = 0; i < n; ++i) g[IDX(i, j, n)] *= f;
    }
}
/* ------------------------------------------------------------------------- */
/*  is_fft products (src/hubbard.cpp:50-81, :159-246): the translation-        */
/*  invariant K is diagonal in momentum space, (inv)expK = F^-1 diag(e^{-+dt   */
/*  E(k)}) F with F the 2-D DFT over the site index i = x + nl*y, applied      */
/*  column by column from the left; a product from the right transposes, so    */
/*  X expK = (expK X^T)^T (expK is symmetric).  The eigenvalues are the DFT of */
/*  K's column 0, E(k) = sum_d K(d, 0) e^{-2 pi i k.d / nl}: the diagonal of F */
/*  K F^-1 that src/hubbard.cpp:63-76 forms with two batches of transforms.    */
/* ------------------------------------------------------------------------- */
void orc_fft_eigens(int nl, double dt, double nnt, double mu, double* lam_f, double* lam_i)
{
    const int n = nl * nl;
    double* K = mat_alloc(n);
    orc_nn_hoppings(nl, K);
    for (size_t e = 0; e < (size_t)n * n; ++e) K[e] *= -nnt;
    for (int i = 0; i < n; ++i) K[IDX(i, i, n)] -= mu;
    for (int ky = 0; ky < nl; ++ky)
        for (int kx = 0; kx < nl; ++kx) {
            double e = 0.0;
            for (int d = 0; d < n; ++d) {
                const int m = (kx * (d % nl) + ky * (d / nl)) % nl;
                e += K[IDX(d, 0, n)] * cos(2.0 * M_PI * m / nl);
            }
            if (lam_f) lam_f[kx + nl * ky] = exp(-dt * e);
            if (lam_i) lam_i[kx + nl * ky] = exp(+dt * e);
        }
    free(K);
}

/* one DFT pass of length nl along x (stride 1) or y (stride nl) of the complex column (re, im) into
 * (ore, oim); sign -1 forward, +1 backward (unscaled) */
static void dft_pass(const orc_ctx* c, const double* re, const double* im, double* ore, double* oim, int along_y,
                     double sign)
{
    const int nl = c->nl;
    for (int a = 0; a < nl; ++a)          /* the other coordinate */
        for (int k = 0; k < nl; ++k) {    /* output frequency (or position) */
            double sr = 0.0, si = 0.0;
            for (int m = 0; m < nl; ++m) {
                const int idx = along_y ? a + nl * m : m + nl * a;
                const int w = (k * m) % nl;
                const double cr = c->tw_c[w], ci = sign * c->tw_s[w];
                sr += re[idx] * cr - im[idx] * ci;
                si += re[idx] * ci + im[idx] * cr;
            }
            const int o = along_y ? a + nl * k : k + nl * a;
            ore[o] = sr;
            oim[o] = si;
        }
}

/* g <- (inv)expK g through the DFT, lam = lam_f or lam_i (fft_mult_from_left, src/hubbard.cpp:172-189) */
static void fft_left(orc_ctx* c, double* g, const double* lam)
{
    const int n = c->n;
    for (int j = 0; j < n; ++j) {
        double* v = g + (size_t)j * n;
        for (int i = 0; i < n; ++i) { c->vre[i] = v[i]; c->vim[i] = 0.0; }
        dft_pass(c, c->vre, c->vim, c->wre, c->wim, 0, -1.0);
        dft_pass(c, c->wre, c->wim, c->vre, c->vim, 1, -1.0);
        for (int k = 0; k < n; ++k) { c->vre[k] *= lam[k]; c->vim[k] *= lam[k]; }
        dft_pass(c, c->vre, c->vim, c->wre, c->wim, 1, +1.0);
        dft_pass(c, c->wre, c->wim, c->vre, c->vim, 0, +1.0);
        for (int i = 0; i < n; ++i) v[i] = c->vre[i] / n;
    }
}
/* g <- g (inv)expK: transpose, from the left, transpose back (src/hubbard.cpp:218-230) */
static void fft_right(orc_ctx* c, double* g, const double* lam)
{
    mat_transpose(c->n, c->t3, g);
    fft_left(c, c->t3, lam);
    mat_transpose(c->n, g, c->t3);
}

/* switch the expK products of a chain to the is_fft form (call before orc_init) */
void orc_set_fft(orc_ctx* c, int on)
{
    c->is_fft = on;
    if (!on || c->lam_f) return;
    const int n = c->n, nl = c->nl;
    c->lam_f = (double*)malloc(sizeof(double) * n); c->lam_i = (double*)malloc(sizeof(double) * n);
    c->tw_c = (double*)malloc(sizeof(double) * nl); c->tw_s = (double*)malloc(sizeof(double) * nl);
    c->wre = (double*)malloc(sizeof(double) * n); c->wim = (double*)malloc(sizeof(double) * n);
    c->vre = (double*)malloc(sizeof(double) * n); c->vim = (double*)malloc(sizeof(double) * n);
    for (int m = 0; m < nl; ++m) { c->tw_c[m] = cos(2.0 * M_PI * m / nl); c->tw_s[m] = sin(2.0 * M_PI * m / nl); }
    orc_fft_eigens(nl, c->dt, c->nnt, c->mu, c->lam_f, c->lam_i);
}

/* expK products of either form: fwd selects expK (else invExpK), left from the left (else right) */
static void expk_mul(orc_ctx* c, double* g, int fwd, int left)
{
    if (c->is_fft) {
        if (left) fft_left(c, g, fwd ? c->lam_f : c->lam_i);
        else fft_right(c, g, fwd ? c->lam_f : c->lam_i);
        return;
    }
    const double* E = fwd ? c->expK : c->invExpK;
    if (left) gemm(c->n, 0, E, 0, g, c->t3);
    else gemm(c->n, 0, g, 0, E, c->t3);
    mat_copy(c->n, g, c->t3);
}

/* B-ops (src/hubbard.cpp:293-341); the expK GEMMs are src/hubbard.cpp:148-156 */
static void mul_B_left(orc_ctx* c, double* g, int t)     { expk_mul(c, g, 1, 1); scale_rows(c, g, t, +1); }
static void mul_B_right(orc_ctx* c, double* g, int t)    { scale_cols(c, g, t, +1); expk_mul(c, g, 1, 0); }
static void mul_invB_left(orc_ctx* c, double* g, int t)  { scale_rows(c, g, t, -1); expk_mul(c, g, 0, 1); }
static void mul_invB_right(orc_ctx* c, double* g, int t) { expk_mul(c, g, 0, 0); scale_cols(c, g, t, -1); }
static void mul_adjB_left(orc_ctx* c, double* g, int t)  { scale_rows(c, g, t, +1); expk_mul(c, g, 1, 1); }

/* exported B-op for unit tests: op 0..4 = B_left, B_right, invB_left, invB_right, adjB_left */
void orc_apply_bop(orc_ctx* c, int op, int t, double* g)
{
    switch (op) {
        case 0: mul_B_left(c, g, t); break;
        case 1: mul_B_right(c, g, t); break;
        case 2: mul_invB_left(c, g, t); break;
        case 3: mul_invB_right(c, g, t); break;
        default: mul_adjB_left(c, g, t); break;
    }
}

/* get_acceptance_ratio (src/hubbard.cpp:123-130) */
static double accept_ratio(orc_ctx* c, int t, int i)
{
    const int n = c->n;
    const double delta = exp(-2.0 * c->alpha * c->fields[t * n + i]) - 1.0;
    const double x = 1.0 + (1.0 - c->gftt[IDX(i, i, n)]) * delta;
    return x * x / (1.0 + delta);
}
/* update_green_function (src/hubbard.cpp:132-139): G -= s * G(:,i) (e_i^T - G(i,:)) */
static void sm_update(orc_ctx* c, int t, int i)
{
    const int n = c->n;
    double* g = c->gftt;
    const double delta = exp(-2.0 * c->alpha * c->fields[t * n + i]) - 1.0;
    const double s = delta / (1.0 + (1.0 - g[IDX(i, i, n)]) * delta);
    double* col = c->col;
    double* row = c->row;
    for (int j = 0; j < n; ++j) { col[j] = s * g[IDX(j, i, n)]; row[j] = (j == i ? 1.0 : 0.0) - g[IDX(i, j, n)]; }
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) g[IDX(j, k, n)] -= col[j] * row[k];
}

/* Core::metropolis_update (src/dqmc_core.cpp:102-119) */
static void metropolis(orc_ctx* c, int t)
{
    const int n = c->n;
    uint32_t* mask = c->accmask + (size_t)t * ((n + 31) / 32);
    memset(mask, 0, sizeof(uint32_t) * ((n + 31) / 32));
    for (int i = 0; i < n; ++i) {
        const double r = accept_ratio(c, t, i);
        const double p = fmin(1.0, fabs(r));
        c->visits++;
        if (rng_bernoulli(&c->rng, p)) {
            sm_update(c, t, i);
            c->fields[t * n + i] = -c->fields[t * n + i];
            c->accepts++;
            mask[i >> 5] |= 1u << (i & 31);
        }
    }
}

/* Core::initialize_svdstacks (:67-87) + initialize_green_functions (:89-97) */
int orc_init(orc_ctx* c)
{
    const int n = c->n;
    int rc = 0;
    c->left.len = 0; c->right.len = 0; c->current_t = 0;
    c->wrap_eq = 0.0; c->wrap_dyn = 0.0;
    mat_eye(n, c->stacks);
    for (int t = c->nt; t >= 1; --t) {
        mul_adjB_left(c, c->stacks, t % c->nt);
        if ((t - 1) % c->pace == 0) { rc |= stack_push(&c->right, c->stacks); mat_eye(n, c->stacks); }
    }
    rc |= equaltime_gf(&c->w, n, &c->left, &c->right, c->gftt);
    return rc;
}

/* Core::sweep_from_0_to_beta (:167-227) */
int orc_sweep_up(orc_ctx* c)
{
    const int n = c->n, nt = c->nt;
    int rc = 0;
    if (c->current_t != 0 || c->left.len != 0) return -10;
    mat_eye(n, c->stacks);
    for (int t = 1; t <= nt; ++t) {
        c->current_t++;
        mul_B_left(c, c->gftt, t % nt);             /* wrap_from_0_to_beta(t-1) (:142-147) */
        mul_invB_right(c, c->gftt, t % nt);
        metropolis(c, t % nt);
        mul_B_left(c, c->stacks, t % nt);
        if (t % c->pace == 0 || t == nt) {
            stack_pop(&c->right);
            rc |= stack_push(&c->left, c->stacks);
            rc |= equaltime_gf(&c->w, n, &c->left, &c->right, c->t1);
            const double e = compare_error(n, c->t1, c->gftt);
            if (e > c->wrap_eq) c->wrap_eq = e;
            mat_copy(n, c->gftt, c->t1);
            mat_eye(n, c->stacks);
        }
        if (!c->thermalization && c->eqtime) mat_copy(n, c->vecgftt + (size_t)(t % nt) * n * n, c->gftt);
    }
    return rc;
}

/* Core::sweep_from_beta_to_0 (:233-284) */
int orc_sweep_down(orc_ctx* c)
{
    const int n = c->n, nt = c->nt;
    int rc = 0;
    if (c->current_t != nt || c->right.len != 0) return -10;
    mat_eye(n, c->stacks);
    for (int t = nt; t >= 1; --t) {
        metropolis(c, t % nt);
        mul_adjB_left(c, c->stacks, t % nt);
        mul_B_right(c, c->gftt, t % nt);             /* wrap_from_beta_to_0(t) (:156-161) */
        mul_invB_left(c, c->gftt, t % nt);
        if ((t - 1) % c->pace == 0) {
            stack_pop(&c->left);
            rc |= stack_push(&c->right, c->stacks);
            rc |= equaltime_gf(&c->w, n, &c->left, &c->right, c->t1);
            const double e = compare_error(n, c->t1, c->gftt);
            if (e > c->wrap_eq) c->wrap_eq = e;
            mat_copy(n, c->gftt, c->t1);
            mat_eye(n, c->stacks);
        }
        if (!c->thermalization && c->eqtime) mat_copy(n, c->vecgftt + (size_t)((t - 1) % nt) * n * n, c->gftt);
        c->current_t--;
    }
    return rc;
}

/* Core::sweep_for_dynamic_green_functions (:291-390) */
int orc_sweep_dynamic(orc_ctx* c)
{
    const int n = c->n, nt = c->nt;
    int rc = 0;
    if (!c->dynamic) return 0;
    if (c->current_t != 0 || c->left.len != 0) return -10;
    mat_copy(n, c->gft0, c->gftt);
    mat_copy(n, c->gf0t, c->gftt);
    for (int i = 0; i < n; ++i) c->gf0t[IDX(i, i, n)] -= 1.0;
    mat_eye(n, c->stacks);
    for (int t = 1; t <= nt; ++t) {
        c->current_t++;
        mul_B_left(c, c->gftt, t % nt);
        mul_invB_right(c, c->gftt, t % nt);
        mul_B_left(c, c->gft0, t % nt);
        mul_invB_right(c, c->gf0t, t % nt);
        mul_B_left(c, c->stacks, t % nt);
        if (t % c->pace == 0 || t == nt) {
            stack_pop(&c->right);
            rc |= stack_push(&c->left, c->stacks);
            double* gtt_new = c->t4;
            rc |= dynamic_gf(&c->w, n, &c->left, &c->right, c->t1, c->t2, gtt_new);
            const double ett = compare_error(n, gtt_new, c->gftt);
            const double et0 = compare_error(n, c->t1, c->gft0);
            const double e0t = compare_error(n, c->t2, c->gf0t);
            if (ett > c->wrap_eq) c->wrap_eq = ett;
            const double ed = et0 > e0t ? et0 : e0t;
            if (ed > c->wrap_dyn) c->wrap_dyn = ed;
            mat_copy(n, c->gftt, gtt_new);
            mat_copy(n, c->gft0, c->t1);
            mat_copy(n, c->gf0t, c->t2);
            mat_eye(n, c->stacks);
        }
        const int ti = t % nt;
        mat_copy(n, c->vecgftt + (size_t)ti * n * n, c->gftt);
        if (ti != 0) {
            mat_copy(n, c->vecgft0 + (size_t)ti * n * n, c->gft0);
            mat_copy(n, c->vecgf0t + (size_t)ti * n * n, c->gf0t);
        } else {
            mat_copy(n, c->vecgft0, c->gftt);
            double* g = c->vecgf0t;
            mat_copy(n, g, c->gftt);
            for (int i = 0; i < n; ++i) g[IDX(i, i, n)] -= 1.0;
        }
    }
    return rc;
}

/* G accessors: kind 0 = current G(t,t), 1 = vecgftt[t], 2 = vecgft0[t], 3 = vecgf0t[t] */
int orc_get_gf(orc_ctx* c, int kind, int t, double* out)
{
    const size_t nn = (size_t)c->n * c->n;
    switch (kind) {
        case 0: memcpy(out, c->gftt, nn * sizeof(double)); return 0;
        case 1: if (!c->vecgftt) return -1; memcpy(out, c->vecgftt + t * nn, nn * sizeof(double)); return 0;
        case 2: if (!c->vecgft0) return -1; memcpy(out, c->vecgft0 + t * nn, nn * sizeof(double)); return 0;
        case 3: if (!c->vecgf0t) return -1; memcpy(out, c->vecgf0t + t * nn, nn * sizeof(double)); return 0;
    }
    return -1;
}
void orc_set_gftt(orc_ctx* c, const double* g) { mat_copy(c->n, c->gftt, g); }
/* get_acceptance_ratio / update_green_function / locally_update_ising_field on the current G */
double orc_acceptance_ratio(orc_ctx* c, int t, int i) { return accept_ratio(c, t, i); }
void orc_update_green_function(orc_ctx* c, int t, int i) { sm_update(c, t, i); }
void orc_flip_field(orc_ctx* c, int t, int i) { c->fields[t * c->n + i] = -c->fields[t * c->n + i]; }
/* matrix_compare_error itself, exported for its own tests */
double orc_compare_error(int n, const double* a, const double* b) { return compare_error(n, a, b); }
void orc_get_stats(orc_ctx* c, double* wrap_eq, double* wrap_dyn, long long* accepts, long long* visits)
{
    *wrap_eq = c->wrap_eq; *wrap_dyn = c->wrap_dyn; *accepts = c->accepts; *visits = c->visits;
}
void orc_get_accmask(orc_ctx* c, uint32_t* out) { memcpy(out, c->accmask, sizeof(uint32_t) * (size_t)c->nt * ((c->n + 31) / 32)); }
int orc_current_t(orc_ctx* c) { return c->current_t; }

/* equal-time G from scratch at slice boundary: exported for direct numerics tests.
 * Builds both stacks from the current fields around time t (0 <= t <= nt, a multiple of pace or nt)
 * and returns compute_equaltime_gf. */
int orc_equaltime_gf_at(orc_ctx* c, int t, double* out)
{
    const int n = c->n, nt = c->nt;
    orc_stack L, R;
    stack_init(&L, n, nt); stack_init(&R, n, nt);
    double* st = mat_alloc(n);
    int rc = 0;
    mat_eye(n, st);
    for (int s = nt; s > t; --s) {   /* right: B(t+1)^T ... B(nt)^T, pushed from the far end */
        mul_adjB_left(c, st, s % nt);
        if ((s - 1) % c->pace == 0 || s - 1 == t) { rc |= stack_push(&R, st); mat_eye(n, st); }
    }
    mat_eye(n, st);
    for (int s = 1; s <= t; ++s) {
        mul_B_left(c, st, s % nt);
        if (s % c->pace == 0 || s == t) { rc |= stack_push(&L, st); mat_eye(n, st); }
    }
    rc |= equaltime_gf(&c->w, n, &L, &R, out);
    stack_free(&L); stack_free(&R); free(st);
    return rc;
}

/* ------------------------------------------------------------------------- */
/*  Observables (include/observable_methods.hpp:39-195), per measurement.      */
/*  Equal-time ones use vecgftt[0..nt-1]; dynamic ones vecgft0 / vecgf0t.      */
/* ------------------------------------------------------------------------- */
void orc_measure_equaltime(orc_ctx* c, double* filling, double* double_occu, double* swave)
{
    const int n = c->n, nt = c->nt;
    double fn = 0.0, docc = 0.0, sw = 0.0;
    for (int t = 0; t < nt; ++t) {
        const double* g = c->vecgftt + (size_t)t * n * n;
        double tr = 0.0;
        for (int i = 0; i < n; ++i) tr += g[IDX(i, i, n)];
        fn += 2.0 - (tr + tr) / n;
        for (int i = 0; i < n; ++i) docc += (1.0 - g[IDX(i, i, n)]) * (1.0 - g[IDX(i, i, n)]);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const double gc = (i == j ? 1.0 : 0.0) - g[IDX(j, i, n)];
                sw += gc * gc + g[IDX(i, j, n)] * g[IDX(i, j, n)];
            }
    }
    *filling = fn / nt;
    *double_occu = docc / ((double)n * nt);
    *swave = sw / ((double)n * nt);
}

/* dynamic: gf[k*nt + t] for the nk momenta given as (kx,ky) pairs; ldos[t]; spin[t] */
void orc_measure_dynamic(orc_ctx* c, int nk, const double* kxy, double* gf, double* ldos, double* spin)
{
    const int n = c->n, nt = c->nt, nl = c->nl;
    for (int t = 0; t < nt; ++t) {
        const double* g = c->vecgft0 + (size_t)t * n * n;
        const double* h = c->vecgf0t + (size_t)t * n * n;
        double tr = 0.0, sc = 0.0;
        for (int i = 0; i < n; ++i) { tr += g[IDX(i, i, n)]; sc -= h[IDX(i, i, n)] * g[IDX(i, i, n)] + h[IDX(i, i, n)] * g[IDX(i, i, n)]; }
        ldos[t] = tr / n;
        spin[t] = sc / (2.0 * n);
        for (int k = 0; k < nk; ++k) {
            const double kx = kxy[2 * k], ky = kxy[2 * k + 1];
            double acc = 0.0;
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    const int xi = i % nl, yi = i / nl, xj = j % nl, yj = j / nl;
                    const int dx = (xj - xi + nl) % nl, dy = (yj - yi + nl) % nl;
                    /* fourier factor exp(-i k.r), r = displacement(i,j) (square_lattice.cpp:63-81, 120-135) */
                    acc += cos(dx * kx + dy * ky) * g[IDX(j, i, n)] / n;
                }
            gf[k * nt + t] = acc;
        }
    }
}

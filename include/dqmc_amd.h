/*
 * dqmc_amd.h -- C ABI of the MI355X (gfx950) DQMC sweep engine.
 *
 * The engine replaces the reference's sweep path: DQMC::Core (include/dqmc_core.h:35-112)
 * driving Model::HubbardAttractiveU (include/hubbard.h:31-111) with Utils::SvdStack and
 * Utils::StableNumerics underneath.  One context holds `n_chains` independent Markov
 * chains (the reference runs one chain per MPI rank, src/main.cpp:111) resident on one GPU.
 *
 * Conventions
 *   - Every call returns 0 on success or a negative DQMC_E* code; no exception crosses
 *     the ABI.  dqmc_last_error() returns the message of the last failure of a context.
 *   - The caller owns all host buffers.  The context owns device memory and one HIP stream.
 *   - Matrices crossing the ABI are N x N FP64, column-major (Eigen's default, the layout
 *     of the reference's Eigen::MatrixXd, include/dqmc_core.h:27).
 *   - Fields cross the ABI as int8 in [t][i] order (nt x ns), the row order of the
 *     reference's `ising.fields` file (include/dqmc_io.hpp:234-255).
 *   - Calls are stream-ordered and synchronous at return unless stated otherwise.
 *   - One host thread per context.
 */
#ifndef DQMC_AMD_H
#define DQMC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dqmc_ctx dqmc_ctx;

/* Mirrors DQMC::Params (include/dqmc_params.hpp:20-45) for the fields the sweep path reads
 * (keys parsed at include/dqmc_parser.hpp:60-96). */
typedef struct dqmc_params_c {
    int nl;                  /* [Lattice.SquareLattice] nl, ns = ng = nl*nl          */
    int nt;                  /* [MonteCarlo] nt                                      */
    double beta;             /* [MonteCarlo] beta, dt = beta / nt                    */
    double nnt;              /* [Model.HubbardAttractiveU] nnt                       */
    double u;                /* [Model.HubbardAttractiveU] u (absolute value of U)   */
    double mu;               /* [Model.HubbardAttractiveU] mu (default of each chain)*/
    int stabilization_pace;  /* [MonteCarlo] stabilization_pace                      */
    int is_fft;              /* [MonteCarlo] is_fft: accepted; both values apply the same
                              * exp(-+dt K), built from the lattice dispersion (the diagonal
                              * form the reference's FFT path computes, src/hubbard.cpp:50-81) */
    int collect_eqtime;      /* keep G(t,t) of every slice (Measurement::Handle::isEqualTime) */
    int collect_dynamic;     /* keep G(t,0), G(0,t) (Measurement::Handle::isDynamic)  */
} dqmc_params_c;

enum { DQMC_FP64 = 0, DQMC_MIXED = 1 };
enum { DQMC_GF_CURRENT = 0, DQMC_GF_TT = 1, DQMC_GF_T0 = 2, DQMC_GF_0T = 3 };
enum {
    DQMC_OK = 0,
    DQMC_EINVAL = -1,     /* bad argument                                   */
    DQMC_EHIP = -2,       /* HIP runtime error                              */
    DQMC_ESTATE = -3,     /* call not allowed in the current sweep state    */
    DQMC_ENUMERIC = -4,   /* NaN/Inf or a zero pivot in a stabilization     */
    DQMC_ENOMEM = -5,
    DQMC_ENODEV = -6      /* no gfx950 device                               */
};

/* Context lifetime.  precision: DQMC_FP64 (everything FP64) or DQMC_MIXED (wraps and the
 * delayed-update flushes in FP32, stabilization / recompose / accept ratios FP64).
 * delay_k (1..32): sites are visited in blocks of delay_k; accept/reject inside a block runs
 * on the block of G, and the block's accepted flips are applied to G as one rank-|A| update.
 * delay_k = 1 applies every accepted flip immediately, the reference's rank-1
 * Sherman-Morrison update (src/hubbard.cpp:132-139).  The accept/reject sequence does not
 * depend on delay_k.  Lattices: 1 <= nl <= 32 (N <= 1024).  Above 16 x 16 the pending update
 * of a block is limited by LDS, so delay_k is capped (FP64: 24 at nl = 18, 12 at nl = 24, 8 at
 * nl = 32); the FP64 slice products there use the ring factor padded to 32 x 32.
 * Replaces: construction of DQMC::Core + Model::HubbardAttractiveU + SquareLattice and their
 * initialize() calls (src/main.cpp:119-128; src/dqmc_core.cpp:20-30; src/hubbard.cpp:25-48). */
int dqmc_create(const dqmc_params_c* params, int n_chains, int precision, int delay_k, dqmc_ctx** out);
void dqmc_destroy(dqmc_ctx* ctx);
/* Message of the last failure of ctx; with ctx == NULL, of the calling thread's last failed
 * call that has no context (dqmc_create before the context exists, dqmc_apply_bop_noalloc). */
const char* dqmc_last_error(dqmc_ctx* ctx);

/* Switch on the storage of G(t,t) / G(t,0), G(0,t) for measurements after creation (the
 * reference decides this in Core::initialize from the measurement handle,
 * src/dqmc_core.cpp:28-29 and :43-65).  Allocates on first use; call before dqmc_init. */
int dqmc_set_collect(dqmc_ctx* ctx, int collect_eqtime, int collect_dynamic);

/* Per-chain chemical potential (the batched workload spreads mu over chains).
 * Must be called before dqmc_init. */
int dqmc_set_chain_mu(dqmc_ctx* ctx, int chain, double mu);

/* Replaces Utils::Random::set_seed (src/random.cpp:15): each chain owns one
 * std::default_random_engine (minstd_rand0) stream, seeded with libstdc++ semantics. */
int dqmc_set_seed(dqmc_ctx* ctx, int chain, uint64_t seed);
int dqmc_get_rng_state(dqmc_ctx* ctx, int chain, uint64_t* state);

/* Replaces Model::HubbardAttractiveU::set_ising_fields_to_random (src/hubbard.cpp:102-110):
 * bernoulli(0.5) -> +1/-1 in t-major order, drawn from the chain's stream. */
int dqmc_fields_random(dqmc_ctx* ctx, int chain);
/* Replaces DQMC::IO::load/save_ising_fields_configs' access to m_ising_fields
 * (include/dqmc_io.hpp:234-290).  s: nt*ns int8 in [t][i] order, values +1/-1. */
int dqmc_set_fields(dqmc_ctx* ctx, int chain, const int8_t* s);
int dqmc_get_fields(dqmc_ctx* ctx, int chain, int8_t* s);

/* Replaces Core::initialize_svdstacks + Core::initialize_green_functions
 * (src/dqmc_core.cpp:67-97) for all chains. */
int dqmc_init(dqmc_ctx* ctx);
/* Replaces Model::HubbardAttractiveU::multiply_B_from_left / multiply_B_from_right /
 * multiply_invB_from_left / multiply_invB_from_right / multiply_adjB_from_left
 * (include/hubbard.h:58-62, src/hubbard.cpp:293-341) for ops 0..4, applied in place to an
 * N x N column-major FP64 matrix with chain `chain`'s fields of slice t (FP64 path). */
int dqmc_apply_bop(dqmc_ctx* ctx, int chain, int op, int t, double* inout);

/* Single-site operations on one chain between sweeps (FP64 contexts), on its current G and fields:
 * dqmc_acceptance_ratio replaces Model::HubbardAttractiveU::get_acceptance_ratio
 * (src/hubbard.cpp:123-130), dqmc_update_green_function replaces update_green_function (:132-139,
 * G for the flip of s(t,i), the field itself unchanged) and dqmc_flip_field replaces
 * locally_update_ising_field (:118-121).  The sweeps apply the same three steps internally. */
int dqmc_acceptance_ratio(dqmc_ctx* ctx, int chain, int t, int i, double* ratio);
int dqmc_update_green_function(dqmc_ctx* ctx, int chain, int t, int i);
int dqmc_flip_field(dqmc_ctx* ctx, int chain, int t, int i);

/* Replaces Core::set_thermalization (include/dqmc_core.h:90). */
int dqmc_set_thermalization(dqmc_ctx* ctx, int is_thermalization);

/* One directional sweep of every chain.
 * Replace Core::sweep_from_0_to_beta (src/dqmc_core.cpp:167-227),
 * Core::sweep_from_beta_to_0 (:233-284) and
 * Core::sweep_for_dynamic_green_functions (:291-390). */
int dqmc_sweep_up(dqmc_ctx* ctx);
int dqmc_sweep_down(dqmc_ctx* ctx);
int dqmc_sweep_dynamic(dqmc_ctx* ctx);

/* Replaces the accessors Core::current_gftt / gftt(t) / gft0(t) / gf0t(t)
 * (include/dqmc_core.h:76-79).  out: N*N FP64 column-major.  For DQMC_MIXED the current
 * G is the FP32 working copy widened to FP64. */
int dqmc_get_gf(dqmc_ctx* ctx, int chain, int kind, int t, double* out);
/* Overwrite a chain's current G(t,t) (FP64 and the FP32 working copy); test hook. */
int dqmc_set_gf(dqmc_ctx* ctx, int chain, const double* g);

/* Replaces Core::equaltimeWrapError / dynamicWrapError (include/dqmc_core.h:80-81); the
 * error metric is the reference's signed max (include/utils/stable_numerics.hpp:42).
 * max_abs_dG: max over the run of max_ij |G_working - G_FP64| at each stabilization, with
 * G_FP64 recomputed from the same fields (the mixed-precision error; equals the absolute
 * wrap error in FP64 mode). */
int dqmc_get_stats(dqmc_ctx* ctx, int chain, double* wrap_eq, double* wrap_dyn, double* max_abs_dG,
                   long long* accepts, long long* visits);
/* Accept bits of the last updating sweep: nt x ceil(ns/32) uint32, bit i%32 of word
 * [t][i/32] set when the flip at (t, i) was accepted (src/dqmc_core.cpp:110). */
int dqmc_get_accept_mask(dqmc_ctx* ctx, int chain, uint32_t* out);

/* Wall-clock timing helpers for benchmarks: record events around the next sweeps on the
 * context's stream and return the elapsed milliseconds between two marks. */
int dqmc_synchronize(dqmc_ctx* ctx);
int dqmc_event_record(dqmc_ctx* ctx, int slot);
int dqmc_event_elapsed(dqmc_ctx* ctx, int slot_begin, int slot_end, float* ms);

/* Per-launch timing of one hot kernel during the next sweeps (benchmark instrumentation):
 * kernel 0 = Metropolis slice update, 1 = Householder QR, 2 = GEMM (all), 3 = triangular solve,
 * 4 = Kronecker-factored exp(-dt K) products (wraps, B-chain).
 * Each launch of that kernel is bracketed by HIP events on the context's stream; end() returns
 * the summed device time and the number of launches.  Eager sweeps only (graphs are bypassed
 * while a timer is active).  No counterpart in the reference. */
int dqmc_kernel_timer_begin(dqmc_ctx* ctx, int kernel, int max_launches);
int dqmc_kernel_timer_end(dqmc_ctx* ctx, double* total_ms, int* launches);

/* Capture each sweep kind into a hipGraph on first use and replay it afterwards
 * (1 = on, default 0). */
int dqmc_use_graphs(dqmc_ctx* ctx, int on);

/* GEMM tile chosen by the autotuner in dqmc_init for FP64 (and FP32) products:
 * 0 = 64x64, 1 = 32x32, 2 = 64x32 output tiles, 3 = 64x64 with double-buffered LDS and global
 * loads two steps ahead, 4 = 128x128 for FP64 products with both output sides >= 128 and an inner
 * dimension >= 64 (tile 3 otherwise); ms (optional, 2 x 5 floats) receives the measured time per wrap-shaped product of
 * every candidate.  No counterpart in the reference. */
int dqmc_get_tuning(dqmc_ctx* ctx, int* tile_fp64, int* tile_fp32, float* ms);
/* Override the autotuner's choice (after dqmc_init, before the first captured sweep). */
int dqmc_set_tuning(dqmc_ctx* ctx, int tile_fp64, int tile_fp32);

/* Fixed-field forward of the differentiable quantities (FP64 contexts).  For every chain, at
 * its current fields and without Metropolis updates:
 *   logdet[b]          = log|det(1 + B_L ... B_1)|  (the fermion determinant of one spin
 *                        species; the reference's weight is its square times the field
 *                        factor e^{-alpha sum s}, src/hubbard.cpp:123-130),
 *   gdiag[b][t][i]     = G(t)(i, i) for the equal-time G of slice t in the storage order of
 *                        Core::m_vecgftt (src/dqmc_core.cpp:370-382).
 * Either output may be NULL.  Leaves the context as dqmc_init does (G recomputed, run
 * statistics unchanged).  No counterpart in the reference. */
int dqmc_forward_fixed(dqmc_ctx* ctx, double* logdet, double* gdiag);

/* Adjoint of dqmc_forward_fixed with respect to (nnt, mu, u) (FP64 contexts):
 *   grad[b][0..2] = gbar_logdet[b] * d logdet[b]/d(nnt, mu, u)
 *                   + sum_{t,i} gbar_gdiag[b][t][i] * d G_b(t)(i,i)/d(nnt, mu, u)
 * at fixed fields (alpha follows u through cosh(alpha) = e^{dt u/2}).  Either cotangent may be
 * NULL (zero).  The derivative of G(t) is contracted through the time-displaced G of all slice
 * pairs, one stabilized dynamic pass per slice (cost ~ L dynamic sweeps).  The d/du component
 * diverges as u -> 0 (alpha ~ sqrt(dt u)).  Leaves the context as dqmc_init does.
 * No counterpart in the reference. */
int dqmc_adjoint(dqmc_ctx* ctx, const double* gbar_logdet, const double* gbar_gdiag, double* grad);

/* Numeric health (no counterpart in the reference, which throws only on SVD non-convergence,
 * include/utils/linear_algebra.hpp:77-80, and otherwise lets NaN propagate).  Every dqmc_init,
 * dqmc_sweep_*, dqmc_forward_fixed and dqmc_adjoint call clears the per-chain flags, runs to
 * completion, and returns DQMC_ENUMERIC when a chain raised one:
 *   DQMC_HEALTH_NONFINITE_G  NaN/Inf in the working G or its FP64 recompute at a stabilization
 *                            (a zero pivot of the recompose solve ends here too)
 *   DQMC_HEALTH_ZERO_PIVOT   zero or non-finite |R_ii| in a UDV factorization (QR breakdown)
 *   DQMC_HEALTH_RATIO        non-finite acceptance ratio (non-finite G_ii at a site visit)
 *   DQMC_HEALTH_WATCHDOG     max_ij |G_recomputed - G_working| at a stabilization above the bound
 *                            set by dqmc_set_watchdog (0 = off, the default)
 * After DQMC_ENUMERIC the flagged chain's G is not trustworthy; dqmc_init rebuilds it from the
 * fields.  dqmc_get_health returns the flags of the last call for one chain. */
enum {
    DQMC_HEALTH_NONFINITE_G = 1,
    DQMC_HEALTH_ZERO_PIVOT = 2,
    DQMC_HEALTH_RATIO = 4,
    DQMC_HEALTH_WATCHDOG = 8
};
int dqmc_set_watchdog(dqmc_ctx* ctx, double max_dG);
int dqmc_get_health(dqmc_ctx* ctx, int chain, uint32_t* flags);

/* Observables on the device, from the stored Green's functions of every slice.
 * Replace Measurement::Handle::initialize / measure_equaltime_observables /
 * measure_dynamic_observables / normalize_cache + push_cache_to_data + clear_cache
 * (src/measure_handle.cpp:18-104) and the estimators of Observable::Methods
 * (include/observable_methods.hpp:39-195), for every chain of the context (each chain keeps its
 * own bins, as each MPI rank of the reference does).
 *   obs_mask: OR of the DQMC_OBS_* bits.  Equal-time observables need dqmc_set_collect(eqtime),
 *             dynamic ones dqmc_set_collect(dynamic), both before dqmc_measure_setup.
 *   nk, kxy:  the momentum list of DynamicGreenFunctions, nk (kx, ky) pairs (the k-stars of
 *             src/square_lattice.cpp:83-95 selected by the momentum_list key).
 * dqmc_measure_equaltime reads G(t,t) of all slices stored by the last updating sweep;
 * dqmc_measure_dynamic reads G(t,0), G(0,t) of the last dynamic sweep.  dqmc_measure_push_bin
 * divides each cache by its sample count, stores it as bin `bin` and clears it (an empty cache is
 * the reference's "cache divided by zero" error).  dqmc_get_bins copies one chain's bins of one
 * observable: [nbins] for the scalars, [nbins][nk][nt] for DQMC_OBS_GREEN_KT, [nbins][nt] for
 * DQMC_OBS_LDOS and DQMC_OBS_SPIN. */
enum {
    DQMC_OBS_FILLING = 1,       /* FillingNumber                         */
    DQMC_OBS_DOUBLE_OCCU = 2,   /* DoubleOccupation                      */
    DQMC_OBS_SWAVE = 4,         /* StaticSWavePairingCorrelation         */
    DQMC_OBS_GREEN_KT = 8,      /* DynamicGreenFunctions                 */
    DQMC_OBS_LDOS = 16,         /* LocalDensityOfStates                  */
    DQMC_OBS_SPIN = 32          /* LocalDynamicTransverseSpinCorrelation */
};
int dqmc_measure_setup(dqmc_ctx* ctx, unsigned obs_mask, int nbins, int nk, const double* kxy);
int dqmc_measure_equaltime(dqmc_ctx* ctx);
int dqmc_measure_dynamic(dqmc_ctx* ctx);
int dqmc_measure_push_bin(dqmc_ctx* ctx, int bin);
int dqmc_get_bins(dqmc_ctx* ctx, int chain, unsigned obs, double* out);

/* Allocation-free B-operations.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).  Applies operation
 * `op` of dqmc_apply_bop (0: B X, 1: X B, 2: B^{-1} X, 3: X B^{-1}, 4: B^T X, with
 * B = diag(e^{dt mu + alpha s}) exp(dt nnt A) of include/hubbard.h:58-62) in place to n_chains
 * N x N column-major FP64 matrices in device memory d_inout (chain c at c * N * N), with the fields
 * of one slice in device memory d_fields (int8 +1/-1, chain c at c * N).  params supplies nl (<= 16),
 * nt, beta, nnt, u and mu (one mu for all chains).  Work is queued on `stream` (a hipStream_t;
 * NULL = the default stream) and the call returns without waiting; it calls no hipMalloc and no
 * allocating routine.  Intermediates live in a library-owned device array of
 * DQMC_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: n_chains * N * N * 8 bytes
 * must fit in it beside 8 KiB of constants, else DQMC_EINVAL.  All calls on one GPU share that
 * array, so calls that overlap on one GPU (different streams or threads) race.  Results equal
 * dqmc_apply_bop's.  Errors are reported by dqmc_last_error(NULL) of the calling thread. */
#define DQMC_NOALLOC_SCRATCH_BYTES (32LL << 20)
int dqmc_apply_bop_noalloc(const dqmc_params_c* params, int n_chains, int op, const int8_t* d_fields,
                           double* d_inout, void* stream);

/* Lattice size / chain count accessors. */
int dqmc_dims(dqmc_ctx* ctx, int* ns, int* nt, int* n_chains);

#ifdef __cplusplus
}
#endif

#endif /* DQMC_AMD_H */

/*
 * fsmc.h - C ABI of the MI355X (gfx950) Monte Carlo pricing engine.
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (FSMC_OK == 0, negative on failure) and never exits the process.  Two
 * flavours exist for each workload:
 *
 *   fsmc_X(ctx, ...)          synchronous; host pointers in and out; the
 *                             context owns and reuses the device buffers.
 *   fsmc_X_async(..., stream) device pointers; the caller provides a device
 *                             workspace of fsmc_X_workspace(...) bytes and a
 *                             hipStream_t (passed as void*); nothing is
 *                             allocated and nothing synchronises.
 *
 * Normals come either from the in-kernel Philox4x32-10 generator
 * (FSMC_RNG_PHILOX) or from a caller buffer laid out exactly as the reference
 * sample consumes its MKL stream (FSMC_RNG_BUFFER, "oracle mode").
 *
 * Philox contract (mirrored bit for bit by the CPU oracle): key =
 * (seed & 0xffffffff, seed >> 32); counter = (c0, c1, c2, tag << 28 ^ stream)
 * with tag 1 GBM, 2 Heston, 3 LIBOR, 4 LSMC.  Each call yields four words
 * (a, b, c, d) -> Box-Muller pairs (a, b) and (c, d): u1 = fl(w1 | 1) 2^-32,
 * u2 = fl(w2) 2^-32 with fl the round-to-nearest-even conversion of the
 * 32-bit word to float, z = sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2).
 *   GBM     c0 = path / 4,  c1 = 0,        c2 = option; normal j -> path 4 c0 + j
 *   Heston  c0 = path,      c1 = step / 2, c2 = option; (z1, z2) of step 2 c1
 *           are normals 0, 1 and of step 2 c1 + 1 normals 2, 3
 *   LIBOR   c0 = path,      c1 = step / 4, c2 = 0;      normal j -> step 4 c1 + j
 *   LSMC    batched (Brownian bridge): c0 = path / 4, c1 = i, c2 = option;
 *           normal j -> Z_i of path 4 c0 + j
 *           reference algorithm: c0 = path, c1 = step / 4, c2 = 0;
 *           normal j -> step 4 c1 + j
 */
#ifndef FSMC_H_
#define FSMC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FSMC_OK 0
#define FSMC_ERR_INVALID_ARG (-1) /* bad size, pointer, mode or parameter   */
#define FSMC_ERR_DEVICE (-2)      /* HIP runtime error                      */
#define FSMC_ERR_ALLOC (-3)       /* device allocation failed               */
#define FSMC_ERR_WORKSPACE (-4)   /* async workspace smaller than required  */

/* ---- mode --------------------------------------------------------------- */
#define FSMC_FP32 0 /* FP32 path arithmetic, compensated FP32 sums, FP64 combine */
#define FSMC_FP64 1 /* plain FP64 throughout                                      */

#define FSMC_RNG_PHILOX 0 /* in-kernel Philox4x32-10 + Box-Muller          */
#define FSMC_RNG_BUFFER 1 /* caller-supplied normals (reference layout)     */
#define FSMC_RNG_MRG32K3A 2 /* the sample's own MKL stream, generated on the GPU
                             * (LIBOR and the LSMC reference algorithm only) */

/* Reference-compatibility switches ("refcompat"): each reproduces one
 * documented behaviour of the reference sample bit for bit. */
#define FSMC_QUIRK_ASIAN_VAR_AS_VOL (1u << 0)   /* MCAsianOptions.cpp:57-60: sigma = v+ */
#define FSMC_QUIRK_ASIAN_RUNNING_VAL2 (1u << 1) /* MCAsianOptions.cpp:191-192,199        */
#define FSMC_QUIRK_LIBOR_INPLACE_ADJ (1u << 2)  /* liborSwaptionGreeks.cpp:173-184       */
#define FSMC_QUIRK_BS_POLY_CND (1u << 3)        /* BlackScholesDP.cpp:84-113: 5-term CND  */

typedef struct fsmc_mode {
  int precision;   /* FSMC_FP32 | FSMC_FP64                               */
  int rng;         /* FSMC_RNG_PHILOX | FSMC_RNG_BUFFER | FSMC_RNG_MRG32K3A */
  uint64_t seed;   /* Philox key                                          */
  uint32_t stream; /* Philox sub-stream id (28 bits), for independent runs */
  uint32_t quirks; /* FSMC_QUIRK_* bit mask                               */
} fsmc_mode;

/* ---- context ------------------------------------------------------------ */
typedef struct fsmc_ctx fsmc_ctx;

int fsmc_ctx_create(int device, fsmc_ctx** out);
int fsmc_ctx_destroy(fsmc_ctx* ctx);
const char* fsmc_status_string(int status);
/* Detail of the last failure on the calling thread ("" if none). */
const char* fsmc_last_error(void);

/* ---- GBM European call ---------------------------------------------------
 * Replaces the option loop inlined in main, MonteCarloInsideBlockingDP.cpp:
 * 237-270 (per option: paths S_T = S*2^(MuByT + VBySqrtT*z), payoff sums
 * v0, v1, then CallResult = e^{-rT} v0/N and CallConfidence = 1.96 * SE,
 * :266-269).  Buffer mode consumes the reference's float samples
 * s = MuByT + VBySqrtT*z (log2 domain, vsRngGaussian output, :249) laid out
 * [n_options][n_paths].  greeks (optional, [n_options][6]) receives the
 * pathwise delta, vega, rho followed by their standard errors.
 * control_variate = 1 (no counterpart in the sample, which uses 0) prices
 * with S_T as control variate: e^{-rT}(mean(y) - beta mean(S_T - S e^{rT})),
 * beta = cov(y, S_T) / var(S_T) from the same paths, y = (S_T - K)^+; the
 * confidence is then 1.96 sqrt(var(y) - beta cov(y, S_T)) / sqrt(N) e^{-rT}.
 * The Greeks are unchanged by it. */
typedef struct fsmc_gbm_params {
  int n_options;
  int n_paths; /* per option */
  double riskfree;
  double volatility;
  int control_variate; /* 0: the sample's estimator; 1: S_T as control variate */
} fsmc_gbm_params;

int fsmc_gbm_european(fsmc_ctx* ctx, const fsmc_gbm_params* p, const fsmc_mode* m, const double* S,
                      const double* X, const double* T, const float* samples, double* call_result,
                      double* call_confidence, double* greeks);
size_t fsmc_gbm_european_workspace(const fsmc_gbm_params* p, const fsmc_mode* m);
int fsmc_gbm_european_async(const fsmc_gbm_params* p, const fsmc_mode* m, const double* d_S,
                            const double* d_X, const double* d_T, const float* d_samples,
                            double* d_call_result, double* d_call_confidence, double* d_greeks,
                            void* d_workspace, size_t workspace_bytes, void* stream);

/* ---- Heston arithmetic-average Asian call --------------------------------
 * Replaces MonteCarloTimeStep, MCAsianOptions.cpp:76-86 (same nine SoA
 * arrays: results, confidences, S, X, T, initialVol, kappa, theta, xi;
 * RISKFREE and RHO of :39-41 become riskfree/rho; SIMSTEPS and RAND_N of
 * :28-33 become n_steps/n_paths).  Buffer mode consumes two normal streams
 * z1, z2 in the reference consumption order
 * [n_options][n_paths/64][n_steps][64] (get_gaussian blocks of 64 lanes x
 * n_steps, mkl_rng.h:57-62, indexed z[step*64+lane] at :151,166); n_paths
 * must then be a multiple of 64.  greeks (optional, [n_options][6], Philox
 * mode without quirks) receives delta dV/dS0, vega dV/dsqrt(v0) and rho dV/dr
 * from the adjoint (reverse-sweep) kernel, followed by their standard
 * errors; the price then comes from the same kernel's forward sweep. */
typedef struct fsmc_heston_params {
  int n_options;
  int n_paths; /* per option */
  int n_steps;
  double riskfree;
  double rho;
} fsmc_heston_params;

int fsmc_heston_asian(fsmc_ctx* ctx, const fsmc_heston_params* p, const fsmc_mode* m, const double* S,
                      const double* X, const double* T, const double* v0, const double* kappa,
                      const double* theta, const double* xi, const double* z1, const double* z2,
                      double* call_result, double* call_confidence, double* greeks);
size_t fsmc_heston_asian_workspace(const fsmc_heston_params* p, const fsmc_mode* m);
int fsmc_heston_asian_async(const fsmc_heston_params* p, const fsmc_mode* m, const double* d_S,
                            const double* d_X, const double* d_T, const double* d_v0,
                            const double* d_kappa, const double* d_theta, const double* d_xi,
                            const double* d_z1, const double* d_z2, double* d_call_result,
                            double* d_call_confidence, double* d_greeks, void* d_workspace,
                            size_t workspace_bytes, void* stream);
/* The same three Greeks by forward (tangent) propagation: each path carries
 * dS/dsqrt(v0) and dv/dsqrt(v0) along its steps (dS/dS0 = S/S0 and
 * dS/dr = S k dt need no recursion), so no checkpoints are stored and no
 * reverse sweep runs; the price is bit for bit the forward kernel's and the
 * Greeks equal the adjoint's to rounding.  No counterpart in the samples:
 * an engine option, cheaper than the adjoint for these three Greeks (the
 * serving batch prices its Heston requests this way).  Philox mode without
 * quirks; d_greeks [n_options][6] as above; the workspace holds the
 * per-workgroup partials (fsmc_heston_asian_tangent_workspace bytes, at
 * most fsmc_heston_asian_workspace's). */
size_t fsmc_heston_asian_tangent_workspace(const fsmc_heston_params* p, const fsmc_mode* m);
int fsmc_heston_asian_tangent_async(const fsmc_heston_params* p, const fsmc_mode* m, const double* d_S,
                                    const double* d_X, const double* d_T, const double* d_v0, const double* d_kappa,
                                    const double* d_theta, const double* d_xi, double* d_result, double* d_conf,
                                    double* d_greeks, void* d_workspace, size_t workspace_bytes, void* stream);
/* Allocation-free variant of fsmc_heston_asian_async.  It has no counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels run with their workspace in a fixed device
 * array of fsmc_heston_asian_noalloc_limit() bytes (16 MiB) that the library
 * holds on each GPU it runs on until the process ends; the call allocates
 * nothing and rejects, with FSMC_ERR_WORKSPACE, a pricing whose workspace
 * (fsmc_heston_asian_workspace without Greeks; with Greeks the adjoint's
 * checkpoints, 16 B per path and 16 steps per option) exceeds it.  The
 * stream must belong to the current device.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (two streams, two threads) race
 * and must be serialised by the caller. */
size_t fsmc_heston_asian_noalloc_limit(void);
int fsmc_heston_asian_noalloc(const fsmc_heston_params* p, const fsmc_mode* m, const double* d_S,
                              const double* d_X, const double* d_T, const double* d_v0, const double* d_kappa,
                              const double* d_theta, const double* d_xi, const double* d_z1, const double* d_z2,
                              double* d_call_result, double* d_call_confidence, double* d_greeks, void* stream);

/* ---- LIBOR market model swaption portfolio + pathwise adjoint deltas -------
 * Replaces the per-path body of liborSwaptionGreeks, liborSwaptionGreeks.cpp:
 * 241-356 (declared liborSwaptionGreeks.hpp:12-13): 80 forward rates, 40
 * log-Euler steps of tenor 0.25 (params.h:27-34), the 15 payer swaptions of
 * params.h:38-40, and the adjoint sweep path_calc_b2.  L0 and lambda are the
 * flat initial rate and volatility of :327-329 (0.051 and 0.2 there).
 * Buffer mode consumes z laid out [n_paths][40] (z[path*40 + n], :386).
 * Outputs: value[2] = {mean discounted portfolio value, its standard error};
 * deltas[160] = {dV/dL0_i for i = 0..79, then their standard errors}.  The
 * means are per path (the sample prints 10x these, NLOOPS = 10, :355-356).
 * Trajectory (the sample stores all 41 x 80 rates of a path, :76,106-107):
 * FP64 stores the 2380 updated rates per path in the workspace; FP32 (lambda
 * <= 0.3) rebuilds row n from row n + 1 in the reverse sweep by inverting the
 * step (truncation <= 1.5e-8 relative), with checkpoint_every = 8 from rows
 * 8, 16, 24 and 32 stored by the forward sweep (each segment recomputed from
 * the exact row at its top; 240 floats per path of workspace), with 0 from
 * the final rates alone. */
typedef struct fsmc_libor_params {
  int n_paths;
  double L0;
  double lambda;
  int checkpoint_every; /* 0 or 8: see above */
} fsmc_libor_params;

/* m->rng = FSMC_RNG_MRG32K3A (synchronous entry point only): z is not read;
 * the normals are generateRand's (mklRng.cpp:44-139): MRG32k3a seeded with
 * m->seed, stream positions k = m->stream * 40 n_paths + 40 path + step. */
int fsmc_libor_swaption_greeks(fsmc_ctx* ctx, const fsmc_libor_params* p, const fsmc_mode* m, const double* z,
                               double* value, double* deltas);
size_t fsmc_libor_workspace(const fsmc_libor_params* p, const fsmc_mode* m);
/* Fills z[n_paths][40] with mean + stddev * (the normals Philox mode uses):
 * a buffer-mode run on this z equals the Philox-mode run. */
int fsmc_libor_normals(fsmc_ctx* ctx, int n_paths, const fsmc_mode* m, double mean, double stddev, double* z);
int fsmc_libor_swaption_greeks_async(const fsmc_libor_params* p, const fsmc_mode* m, const double* d_z,
                                     double* d_value, double* d_deltas, void* d_workspace,
                                     size_t workspace_bytes, void* stream);

/* ---- MKL-compatible MRG32k3a stream -----------------------------------------
 * The stream generateRand draws (LiborMarketModel/mklRng.h:9,
 * mklRng.cpp:44-139: vslNewStream(VSL_BRNG_MRG32K3A, seed), per-thread
 * vslSkipAheadStream, vdRngGaussian/vsRngGaussian with
 * VSL_RNG_METHOD_GAUSSIAN_ICDF), generated on the GPU.  Position k of the
 * stream is written to out[k - offset] for offset <= k < offset + n.
 * Uniforms equal MKL's bit for bit: ((x_k - y_k) mod m1) * (1 / m1).  Normals
 * are mean + stddev * icdf(u_k) with an FP64 inverse normal CDF accurate to
 * ~1e-16 relative; MKL's own ICDF differs from the exact inverse by up to
 * ~7.3e-9 relative, so the normals match MKL's to that. */
int fsmc_mrg32k3a_uniform(fsmc_ctx* ctx, uint32_t seed, uint64_t offset, long long n, double* u);
int fsmc_mrg32k3a_gaussian(fsmc_ctx* ctx, uint32_t seed, uint64_t offset, long long n, double mean,
                           double stddev, double* z);
int fsmc_mrg32k3a_gaussian_async(uint32_t seed, uint64_t offset, long long n, double mean, double stddev,
                                 double* d_z, void* stream);

/* ---- MKL-compatible MT19937 stream ------------------------------------------
 * The generator of the AsianOptionsPricing sample (mkl_rng.h:5-69: an
 * mkl_rng object owns vslNewStream(VSL_BRNG_MT19937, seed) and get_gaussian
 * calls vdRngGaussian(VSL_RNG_METHOD_GAUSSIAN_BOXMULLER2, ..., 0, 1);
 * MCAsianOptions.cpp:92-93 makes two, seeds 42 and 43), on the GPU.  A
 * generator is one sequential stream: every call continues where the last
 * one stopped, as successive vdRngGaussian calls on one VSL stream do, so
 * calls must be ordered (one stream, or synchronised).  The 32-bit outputs
 * equal MKL's bit for bit (init_by_array({seed}) seeding).  Normals come in
 * Box-Muller pairs from consecutive outputs a, b (n even): u1 = a 2^-32 (a = 0
 * taken as 2^-33), u2 = b 2^-32, z = mean + stddev sqrt(-2 ln u1) (sin 2 pi
 * u2, cos 2 pi u2), in FP64; MKL's own Box-Muller arithmetic is not
 * reproducible exactly (its normals differ from this formula by ~2e-8). */
typedef struct fsmc_mt19937 fsmc_mt19937;
int fsmc_mt19937_create(fsmc_ctx* ctx, uint32_t seed, fsmc_mt19937** out);
int fsmc_mt19937_bits(fsmc_mt19937* g, long long n, uint32_t* out);
int fsmc_mt19937_gaussian(fsmc_mt19937* g, long long n, double mean, double stddev, double* out);
int fsmc_mt19937_bits_async(fsmc_mt19937* g, long long n, uint32_t* d_out, void* stream);
int fsmc_mt19937_gaussian_async(fsmc_mt19937* g, long long n, double mean, double stddev, double* d_out,
                                void* stream);
int fsmc_mt19937_destroy(fsmc_mt19937* g);

/* Jump-ahead polynomial of the MT19937 (generator 0: one step = one 32-bit
 * word) or SFMT19937 (generator 1: one step = one 128-bit word, four
 * outputs) stream: coeffs[0..311] receives x^steps mod phi, phi the minimal
 * polynomial of the step (found by Berlekamp-Massey on the generator's own
 * output bits), coefficient i in bit i % 64 of word i / 64.  The state
 * `steps` words ahead is the XOR, over the set coefficients i, of the state
 * window i words ahead (Haramoto et al., INFORMS J. Computing 20(3), 2008);
 * the generators' skip-ahead uses this.  Host only; returns the degree of
 * phi (19937 or 19968) or a negative status.  No counterpart in the
 * samples, which cannot skip these streams. */
int fsmc_jump_poly(int generator, unsigned long long steps, uint64_t* coeffs, int n_words);

/* The European sample's streams (MonteCarloInsideBlockingDP.cpp:229,249:
 * vslNewStream(VSL_BRNG_SFMT19937, 777 + thread) and vsRngGaussian with
 * VSL_RNG_METHOD_GAUSSIAN_ICDF): SFMT19937 seeded by init_by_array({seed})
 * as MKL does, u = int32(output) 2^-32 + 1/2, z = AS241 inverse CDF (FP64),
 * written as float.  n_streams streams seeded seed0, seed0 + 1, ...; stream
 * t writes its first n_per_stream values to d_out[t * n_per_stream + i].
 * With d_mean / d_stddev (device, one pair per segment of seg values of the
 * whole output) value g is d_mean[g / seg] + d_stddev[g / seg] z: the
 * sample's per-option MuByT + VBySqrtT z with seg = paths per option.
 * Raw outputs equal MKL's bit for bit; MKL's float ICDF is within 2e-5 of
 * the exact inverse used here.  Each stream is sequential (one workgroup,
 * seeded in place); nothing is allocated. */
int fsmc_sfmt19937_gaussian_async(uint32_t seed0, int n_streams, long long n_per_stream, long long seg,
                                  const double* d_mean, const double* d_stddev, float* d_out, void* stream);
/* The same outputs with each stream split into up to 1024 substreams that
 * start by a skip-ahead (fsmc_jump_poly), so that one stream spreads over
 * many workgroups; d_workspace holds fsmc_sfmt19937_workspace(n_streams,
 * n_per_stream) bytes (0: the stream is short and runs sequentially). */
size_t fsmc_sfmt19937_workspace(int n_streams, long long n_per_stream);
int fsmc_sfmt19937_gaussian_skip_async(uint32_t seed0, int n_streams, long long n_per_stream, long long seg,
                                       const double* d_mean, const double* d_stddev, float* d_out, void* d_workspace,
                                       size_t workspace_bytes, void* stream);
/* One stream's first n values, mean + stddev z, to host memory. */
int fsmc_sfmt19937_gaussian(fsmc_ctx* ctx, uint32_t seed, long long n, double mean, double stddev, float* out);

/* ---- Longstaff-Schwartz American puts -------------------------------------
 * Batched LSMC for many options with uneven step and path counts: one
 * exercise date per time step, cash flows discounted at r, regression of the
 * discounted cash flow on {1, x, x^2, x^3} of the in-the-money paths with x an
 * affine rescaling of S/K (moment sums per thread and per wave in the working
 * precision, across waves and workgroups in FP64, solved per date by FP64
 * Cholesky).  Paths are built backwards by the Brownian bridge from
 * Philox normals: Z_i of path p = component p & 3 of counter (p/4, i, option)
 * (tag 4), W_N = sqrt(T) Z_{N-1}, W_k = k/(k+1) W_{k+1} + sqrt(dt k/(k+1)) Z_{k-1},
 * S_k = S0 exp((r - sigma^2/2) k dt + sigma W_k).  It replaces the sweep inlined in
 * MonteCarloAmericanOptions/main.cpp:131-279 (with perform_sims,
 * utilities.cpp:42-88) for the configurations that sample cannot express
 * (rates, many options).  price[n][2] = {price, standard error}; european
 * (optional, [n]) = the discounted European put on the same paths.
 * A plan uploads the option table once and owns its device workspace; run it
 * repeatedly on a stream (it enqueues only kernels), then read results. */
typedef struct fsmc_lsmc_plan fsmc_lsmc_plan;

int fsmc_lsmc_plan_create(fsmc_ctx* ctx, int n_options, const fsmc_mode* m, const double* S0, const double* K,
                          const double* sigma, const double* r, const double* T, const int* steps,
                          const int* paths, fsmc_lsmc_plan** out);
/* Sweeps of a plan (no counterpart in the sample, whose sweep is one loop over
 * a stored path matrix, MonteCarloAmericanOptions/main.cpp:131-266):
 *   FSMC_LSMC_SWEEP_RESIDENT: the whole backward sweep of every option is one
 *     cooperative launch; an option's paths are cut into parts of 16384
 *     (FP32) or 8192 (FP64) paths, one workgroup of a CU each, whose bridge
 *     state and cash flows stay in registers; the parts of an option meet once per date
 *     through device-memory records (no grid-wide barrier).  Needs every
 *     option's parts to fit the resident grid at once;
 *   FSMC_LSMC_SWEEP_STREAMED: one launch per date, state streamed through
 *     device memory (any sizes);
 *   FSMC_LSMC_SWEEP_AUTO: resident when it fits, else streamed.
 * Both draw the same normals, so they price the same paths; the moment sums
 * are combined in different orders (FP64 results agree to rounding).
 * fsmc_lsmc_plan_create uses AUTO; fsmc_lsmc_plan_sweep returns the sweep a
 * plan runs (RESIDENT or STREAMED). */
#define FSMC_LSMC_SWEEP_AUTO 0
#define FSMC_LSMC_SWEEP_RESIDENT 1
#define FSMC_LSMC_SWEEP_STREAMED 2
int fsmc_lsmc_plan_create_sweep(fsmc_ctx* ctx, int n_options, const fsmc_mode* m, const double* S0,
                                const double* K, const double* sigma, const double* r, const double* T,
                                const int* steps, const int* paths, int sweep, fsmc_lsmc_plan** out);
int fsmc_lsmc_plan_sweep(const fsmc_lsmc_plan* plan);
size_t fsmc_lsmc_plan_workspace(const fsmc_lsmc_plan* plan);
int fsmc_lsmc_plan_run(fsmc_lsmc_plan* plan, void* stream);
int fsmc_lsmc_plan_results(fsmc_lsmc_plan* plan, void* stream, double* price, double* european);
int fsmc_lsmc_plan_destroy(fsmc_lsmc_plan* plan);
int fsmc_american_lsmc(fsmc_ctx* ctx, int n_options, const fsmc_mode* m, const double* S0, const double* K,
                       const double* sigma, const double* r, const double* T, const int* steps, const int* paths,
                       double* price, double* european);

/* The sample's own algorithm for one option (MonteCarloAmericanOptions/
 * main.cpp:41-296, utilities.cpp:42-123): GBM without a rate in the log2
 * domain, quadratic regression with ridge 0.01 and a cofactor 3x3 inverse,
 * the betas of step s applied at step s-1, the terminal cash flow taken at
 * step n_steps-1 and scaled by 1/N.  Buffer mode consumes
 * z[(s-1)*n_paths + p] for step s = 1..n_steps of path p.  Outputs the
 * sample's "American" and "European" prices (main.cpp:294-295). */
typedef struct fsmc_lsmc_ref_params {
  int n_paths;
  int n_steps;
  double spot, strike, vol, maturity;
} fsmc_lsmc_ref_params;

/* m->rng = FSMC_RNG_MRG32K3A: z is not read; step s of path p takes the
 * sample's single-thread draw (main.cpp:93-106, utilities.cpp:52-72: stream
 * s-1 of MRG32k3a seed m->seed skipped ahead by nb (s-1), nb = n_paths / 1024,
 * and advanced by n_paths per pricing): stream position
 * nb (s-1) + p + m->stream * n_paths, where m->stream counts previous
 * pricings (the sample's loop index). */
int fsmc_american_lsmc_ref(fsmc_ctx* ctx, const fsmc_lsmc_ref_params* p, const fsmc_mode* m, const double* z,
                           double* american, double* european);
/* The path matrix of the same algorithm, perform_sims (utilities.cpp:42-88,
 * declared utilities.h:34): a[s][p] = a[s-1][p] 2^(fwd_log2 + vsdt_log2 z)
 * from a[0][p] = spot, computed on the GPU and copied to the caller's host
 * rows: rows[s], s = 1..n_steps, each n_paths REALs (double for FSMC_FP64,
 * float for FSMC_FP32); rows[0] is neither read nor written.  The normals
 * as fsmc_american_lsmc_ref takes them (buffer mode: z[(s-1)*n_paths + p]). */
int fsmc_american_paths_ref(fsmc_ctx* ctx, int n_paths, int n_steps, double spot, double fwd_log2, double vsdt_log2,
                            const fsmc_mode* m, const double* z, void* const* rows);

/* ---- Closed form and lattice ---------------------------------------------------
 * Black-Scholes calls and puts (BlackScholes/BlackScholesDP.cpp:177-201) for
 * n options with per-option S, X, T, sigma and one rate r: call[n], put[n]
 * (may be NULL), greeks[n][3] = call delta, vega, rho (may be NULL).  The
 * normal CDF is exact (erfc) unless m->quirks has FSMC_QUIRK_BS_POLY_CND, the
 * sample's Abramowitz-Stegun 26.2.17 polynomial (error below 7.5e-8). */
int fsmc_black_scholes(fsmc_ctx* ctx, int n, const fsmc_mode* m, const double* S, const double* X, const double* T,
                       const double* sigma, double r, double* call, double* put, double* greeks);
int fsmc_black_scholes_async(int n, const fsmc_mode* m, const double* d_S, const double* d_X, const double* d_T,
                             const double* d_sigma, double r, double* d_call, double* d_put, double* d_greeks,
                             void* stream);

/* Cox-Ross-Rubinstein lattice with `steps` steps (1..8000) for n options:
 * the Binomial sample's European call (BinomialOptions/binomial_cpu.cpp:62-101)
 * and, with the flags, puts and American exercise.  price[n]. */
#define FSMC_LATTICE_AMERICAN 1u
#define FSMC_LATTICE_PUT 2u
int fsmc_lattice(fsmc_ctx* ctx, int n, const fsmc_mode* m, const double* S, const double* X, const double* T,
                 const double* sigma, double r, int steps, unsigned flags, double* price);
int fsmc_lattice_async(int n, const fsmc_mode* m, const double* d_S, const double* d_X, const double* d_T,
                       const double* d_sigma, double r, int steps, unsigned flags, double* d_price, void* stream);

/* ---- Mixed request batches (serving) ----------------------------------------
 * No counterpart in the samples, which price one workload per process: many
 * GBM European and Heston-Asian requests of uneven sizes, each returning the
 * price and delta, vega, rho (GBM vega = dV/dsigma; Heston vega =
 * dV/dsqrt(v0)) with standard errors.  Requests are served in micro-batches
 * of requests_per_launch (0 = all) in arrival order, each as one GBM grid, one
 * Heston grid and one finalize grid over (request, chunk) work items.  A
 * request's normals are those of the standalone entry point with the request
 * index as the option index, so a batch reproduces fsmc_gbm_european /
 * fsmc_heston_asian (with Greeks) up to summation order.  latency_us is the
 * time from the start of the run to the end of the request's micro-batch
 * (HIP events).  Runs of one batch share its workspace: do not overlap them. */
#define FSMC_REQ_GBM_EUROPEAN 0
#define FSMC_REQ_HESTON_ASIAN 1

typedef struct fsmc_request {
  int kind;    /* FSMC_REQ_*                 */
  int n_paths;
  int n_steps; /* Heston monitoring steps    */
  double S0, K, T, r;
  double sigma;                      /* GBM     */
  double v0, kappa, theta, xi, rho;  /* Heston  */
} fsmc_request;

typedef struct fsmc_result {
  double price, delta, vega, rho;
  double price_se, delta_se, vega_se, rho_se;
  double latency_us;
} fsmc_result;

typedef struct fsmc_batch_tuning {
  int chunk_paths;         /* paths per work item, a multiple of 1024 */
  int requests_per_launch; /* micro-batch size, 0 = the whole batch   */
  int replay_graph;        /* 1: the batch's launch sequence (kernels and
                              completion stamps) is captured once into a
                              hipGraph that fsmc_batch_run replays        */
  int workgroups_per_cu;   /* 0: one grid of GBM and one of Heston items per
                              micro-batch; 1..16: one persistent grid of
                              CUs x this many workgroups per micro-batch,
                              each running a static longest-first list of
                              items of both kinds                         */
  int work_queue;          /* with workgroups_per_cu >= 1: 0 the static
                              lists above; 1 the persistent workgroups take
                              items from a shared counter, longest first
                              (one relaxed atomic increment per item; no
                              workgroup waits on another)                 */
} fsmc_batch_tuning;

typedef struct fsmc_batch fsmc_batch;

/* tuning NULL: the autotuner's cached choice for a batch of this shape, or
 * {4096, 0, 0, 0, 0}. */
int fsmc_batch_create(fsmc_ctx* ctx, const fsmc_request* req, int n, const fsmc_mode* m,
                      const fsmc_batch_tuning* tuning, fsmc_batch** out);
int fsmc_batch_run(fsmc_batch* b, void* stream);
/* Makes its stream wait for the last fsmc_batch_run (whichever stream that
 * run was enqueued on), copies the results and synchronises its stream. */
int fsmc_batch_results(fsmc_batch* b, void* stream, fsmc_result* out);
/* Times the batch over chunk sizes x micro-batch sizes (grouped grids), then
 * the winner as a persistent grid of 1, 2, 3, 4, 6 or 8 workgroups per CU
 * with static lists and with the work queue, then the best with graph
 * replay; keeps the fastest and caches it for batches of the same shape in
 * this process. */
int fsmc_batch_autotune(fsmc_batch* b, fsmc_batch_tuning* chosen);
/* The same search with the micro-batch size fixed at requests_per_launch
 * (> 0): for latency-bound serving, where each launch's requests complete
 * together and the micro-batch size is set by the caller's latency target.
 * Configures this batch only; the per-shape cache keeps the throughput
 * choice.  No counterpart in the samples (their drivers price one option per
 * call). */
int fsmc_batch_autotune_microbatch(fsmc_batch* b, int requests_per_launch, fsmc_batch_tuning* chosen);
int fsmc_batch_tuning_get(const fsmc_batch* b, fsmc_batch_tuning* out);
int fsmc_batch_destroy(fsmc_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* FSMC_H_ */

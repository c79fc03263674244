/*
 * CPU oracle of the Monte Carlo engine.  TEST INFRASTRUCTURE ONLY: the GPU
 * results are compared against it (and bench.py times it as the CPU
 * baseline); no product code path links or calls it.
 *
 * Every function restates the reference algorithm in plain FP64 C, citing
 * the reference file:line it follows.  The Philox/Box-Muller contract is the
 * one documented in include/fsmc.h.
 */
#ifndef FSMC_ORACLE_H_
#define FSMC_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generator ---------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_box_muller(uint32_t a, uint32_t b, double* z0, double* z1);
/* Four normals of the counter (c0, c1, c2, tag<<28 ^ stream) under key seed. */
void orc_normals4(uint64_t seed, uint32_t stream, uint32_t tag, uint32_t c0, uint32_t c1, uint32_t c2,
                  double out[4]);

/* ---- analytic pins -------------------------------------------------------- */
double orc_bs_call(double S, double X, double T, double r, double sigma);
/* delta, vega, rho of the Black-Scholes call. */
void orc_bs_call_greeks(double S, double X, double T, double r, double sigma, double out[3]);
double orc_crr_european_call(double S, double X, double T, double r, double sigma, int steps);
double orc_crr_american_put(double S, double K, double T, double r, double sigma, int steps);
/* Bermudan put on the same lattice: exercise only at t = k T / dates,
 * k = 0..dates, with steps_per_date lattice steps between dates. */
double orc_crr_bermudan_put(double S, double K, double T, double r, double sigma, int dates, int steps_per_date);
/* Geometric-average Asian call with discrete monitoring t_k = k T/N, k=1..N,
 * under GBM with volatility sigma (closed form; the xi = 0, v0 = theta
 * limit of the Heston model with sigma = sqrt(v0)). */
double orc_geometric_asian_call(double S, double X, double T, double r, double sigma, int N);

/* ---- GBM European ----------------------------------------------------------- */
void orc_gbm_buffer(int n_opt, int n_paths, double r, double sigma, const double* S, const double* X,
                    const double* T, const float* samples, double* res, double* conf);
void orc_gbm_philox(int n_opt, int n_paths, double r, double sigma, const double* S, const double* X,
                    const double* T, uint64_t seed, uint32_t stream, double* res, double* conf,
                    double* greeks6);
/* The same paths priced with S_T as control variate (an engine option, no
 * counterpart in the samples). */
void orc_gbm_buffer_cv(int n_opt, int n_paths, double r, const double* S, const double* X, const double* T,
                       const float* samples, double* res, double* conf);
void orc_gbm_philox_cv(int n_opt, int n_paths, double r, double sigma, const double* S, const double* X,
                       const double* T, uint64_t seed, uint32_t stream, double* res, double* conf);

/* ---- Heston Asian --------------------------------------------------------- */
#define ORC_Q_VAR_AS_VOL 1u
#define ORC_Q_RUNNING_VAL2 2u
void orc_heston_buffer(int n_opt, int n_paths, int n_steps, double r, double rho, const double* S,
                       const double* X, const double* T, const double* v0, const double* kappa,
                       const double* theta, const double* xi, const double* z1, const double* z2,
                       uint32_t quirks, double* res, double* conf);
void orc_heston_philox(int n_opt, int n_paths, int n_steps, double r, double rho, const double* S,
                       const double* X, const double* T, const double* v0, const double* kappa,
                       const double* theta, const double* xi, uint64_t seed, uint32_t stream,
                       uint32_t quirks, double* res, double* conf, double* greeks6);

/* ---- LIBOR market model ------------------------------------------------- */
#define ORC_Q_LIBOR_INPLACE 4u
/* One path: forward rates from the initial curve L0[80] and z[40]; returns
 * the discounted portfolio value and writes dV/dL0 into Lb[80]. */
double orc_libor_path(const double* z, const double* L0, double lambda, uint32_t quirks, double* Lb);
/* value[2] = {mean, se}; deltas[160] = {means, ses}.  z: [n_paths][40]
 * (buffer) or NULL for the engine's Philox normals. */
void orc_libor(int n_paths, double L0, double lambda, const double* z, uint64_t seed, uint32_t stream,
               uint32_t quirks, double* value, double* deltas);

/* ---- Longstaff-Schwartz American put ------------------------------------ */
/* Reference algorithm (refcompat): one option, z[n_steps][n_paths] normals,
 * quadratic basis, ridge 0.01, cofactor inverse, lagged betas, no rate.
 * betas (NULL or [n_steps][3]): row s = the coefficients regressed at date
 * s = n_steps-1 .. 1 (the ones the decisions at date s-1 use). */
void orc_lsmc_ref(int n_paths, int n_steps, double spot, double strike, double vol, double maturity,
                  const double* z, double* american, double* european, double* betas);
/* Correct LSMC: cash flows discounted to 0, cubic basis in the standardized
 * moneyness x = 1 + (S/K - 1)/(sigma sqrt(t)), FP64 Cholesky per date.
 * price2 = [n][2] {price, se}; european may be NULL. */
void orc_lsmc(int n_opt, const double* S0, const double* K, const double* sigma, const double* r,
              const double* T, const int* steps, const int* paths, uint64_t seed, uint32_t stream,
              double* price2, double* european);
/* 4x4 SPD solve used by both the oracle and the kernels; returns 0 on a
 * non-positive pivot. */
int orc_chol4_solve(const double* G, const double* rhs, double* beta);

/* ---- MKL-compatible MRG32k3a stream (the generateRand seam) ---------------
 * L'Ecuyer's MRG32k3a as VSL_BRNG_MRG32K3A seeds it from one 32-bit seed
 * (x = (seed mod m1, 1, 1), y = (1, 1, 1)); uniform i is
 * ((x_i - y_i) mod m1) * (1/m1).  u[k] = uniform offset + k, k < n: the
 * stream order in which generateRand (LiborMarketModel/mklRng.cpp:44-139)
 * fills z, whatever the thread count (each thread skips ahead to its
 * start).  z[k] = mean + stddev * icdf(u[k]) with icdf = Wichura's AS241
 * (PPND16), the inverse normal CDF to ~1e-16 relative (MKL's own ICDF is not
 * reproducible bit for bit and differs from the exact inverse by up to a few
 * 1e-9). */
void orc_mrg32k3a_uniform(uint32_t seed, uint64_t offset, int64_t n, double* u);
double orc_icdf(double u);
void orc_mrg32k3a_gaussian(uint32_t seed, uint64_t offset, int64_t n, double mean, double stddev, double* z);
void orc_mt19937_bits(uint32_t seed, uint64_t offset, int64_t n, uint32_t* out);
void orc_mt19937_gaussian(uint32_t seed, uint64_t offset, int64_t n, double mean, double stddev, double* z);

/* Threads of the OpenMP loops over options; returns the previous setting. */
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif

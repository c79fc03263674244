/*
 * CPU oracle of the Monte Carlo engine -- TEST INFRASTRUCTURE ONLY.
 *
 * Used by tests/ as the results the GPU kernels are compared with and by
 * bench.py as the timed CPU baseline.  Plain FP64 C; each function cites the reference
 * file:line whose algorithm it restates.  Parity pins of this file (see the
 * tests named in each comment):
 *   philox        Random123 known-answer vectors        tests/test_oracle_pins.py
 *   box_muller    moment and identity checks           tests/test_oracle_pins.py
 *   bs/crr        textbook values, convergence, parity tests/test_oracle_pins.py
 *   gbm_*         Black-Scholes price/Greeks within 3 SE
 *   gbm_*_cv      small-K closed form (exact, zero variance), Black-Scholes
 *                 within 4 SE, variance below the plain estimator's
 *   heston_*      exact martingale mean, N=1 reduction to Black-Scholes,
 *                 N=2 conditional-closed-form quadrature (rho, xi, kappa,
 *                 theta, truncation, both vol conventions)
 *   mt19937       MKL's outputs for seed 42 and numpy's RandomState([seed])
 *                 raw draws (the same init_by_array seeding)
 *   heston Q5     closed form of the running-val2 confidence (identical
 *                 deterministic paths)
 *   geometric     exact-GBM Monte Carlo for N = 4 and 12 dates
 *   libor         adjoint = central differences; lambda = 0 closed form;
 *                 deep-in-the-money book = the linear (martingale) value for
 *                 lambda = 0.1, 0.2 (pins the spot-measure drift)
 *   lsmc_ref      numSteps = 2 closed form (1/N terminal scaling, terminal
 *                 date); a hand-placed 5-path case whose decisions need the
 *                 lagged, ridged fit
 * tests/test_oracle_mutants.py plants the mistakes these pins are meant to
 * catch in copies of this file and shows that each pin fails on them.
 */
#include "fsmc_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_LOG2E 1.4426950408889634074
#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, Salmon et al. SC'11: ten rounds of two 32x32->64 products
 * with the Weyl key schedule. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void orc_box_muller(uint32_t a, uint32_t b, double* z0, double* z1) {
  /* fl(): round-to-nearest-even conversion of the 32-bit word to float */
  const double u1 = (double)(float)(a | 1u) * 0x1.0p-32;
  const double u2 = (double)(float)b * 0x1.0p-32;
  const double r = sqrt(-2.0 * log(u1));
  *z0 = r * cos(2.0 * ORC_PI * u2);
  *z1 = r * sin(2.0 * ORC_PI * u2);
}

void orc_normals4(uint64_t seed, uint32_t stream, uint32_t tag, uint32_t c0, uint32_t c1, uint32_t c2,
                  double out[4]) {
  const uint32_t ctr[4] = {c0, c1, c2, (tag << 28) ^ (stream & 0x0FFFFFFFu)};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  orc_box_muller(w[0], w[1], &out[0], &out[1]);
  orc_box_muller(w[2], w[3], &out[2], &out[3]);
}

/* ------------------------------------------------------------------------ */
/* Black-Scholes call, MonteCarloInsideBlockingDP.cpp:79-93 (erf form). */
static double ncdf(double x) { return 0.5 + 0.5 * erf(x / sqrt(2.0)); }
static double npdf(double x) { return exp(-0.5 * x * x) / sqrt(2.0 * ORC_PI); }

double orc_bs_call(double S, double X, double T, double r, double sigma) {
  const double sq = sigma * sqrt(T);
  const double d1 = (log(S / X) + (r + 0.5 * sigma * sigma) * T) / sq;
  const double d2 = d1 - sq;
  return S * ncdf(d1) - X * exp(-r * T) * ncdf(d2);
}

void orc_bs_call_greeks(double S, double X, double T, double r, double sigma, double out[3]) {
  const double sq = sigma * sqrt(T);
  const double d1 = (log(S / X) + (r + 0.5 * sigma * sigma) * T) / sq;
  const double d2 = d1 - sq;
  out[0] = ncdf(d1);
  out[1] = S * npdf(d1) * sqrt(T);
  out[2] = X * T * exp(-r * T) * ncdf(d2);
}

/* Cox-Ross-Rubinstein lattice, binomial_cpu.cpp:62-101: u = e^{sigma sqrt(dt)},
 * d = 1/u, pu = (e^{r dt} - d)/(u - d), backward induction with the
 * discounted probabilities.  r and sigma are exact doubles here (the sample
 * promotes float constants, binomial.h:35-36). */
double orc_crr_european_call(double S, double X, double T, double r, double sigma, int steps) {
  const double dt = T / steps;
  const double vdt = sigma * sqrt(dt);
  const double If = exp(r * dt), Df = exp(-r * dt);
  const double u = exp(vdt), d = exp(-vdt);
  const double pu = (If - d) / (u - d), pd = 1.0 - pu;
  const double puDf = pu * Df, pdDf = pd * Df;
  double* call = (double*)malloc(sizeof(double) * (steps + 1));
  for (int i = 0; i <= steps; ++i) {
    const double v = S * exp(vdt * (2.0 * i - steps)) - X;
    call[i] = v > 0.0 ? v : 0.0;
  }
  for (int i = steps; i > 0; --i)
    for (int j = 0; j <= i - 1; ++j) call[j] = puDf * call[j + 1] + pdDf * call[j];
  const double res = call[0];
  free(call);
  return res;
}

/* Same lattice, American put: each node takes max(continuation, exercise). */
double orc_crr_american_put(double S, double K, double T, double r, double sigma, int steps) {
  const double dt = T / steps;
  const double vdt = sigma * sqrt(dt);
  const double If = exp(r * dt), Df = exp(-r * dt);
  const double u = exp(vdt), d = exp(-vdt);
  const double pu = (If - d) / (u - d), pd = 1.0 - pu;
  const double puDf = pu * Df, pdDf = pd * Df;
  double* put = (double*)malloc(sizeof(double) * (steps + 1));
  for (int i = 0; i <= steps; ++i) {
    const double v = K - S * exp(vdt * (2.0 * i - steps));
    put[i] = v > 0.0 ? v : 0.0;
  }
  for (int i = steps; i > 0; --i) {
    for (int j = 0; j <= i - 1; ++j) {
      const double cont = puDf * put[j + 1] + pdDf * put[j];
      const double ex = K - S * exp(vdt * (2.0 * j - (i - 1)));
      put[j] = cont > ex ? cont : ex;
    }
  }
  const double res = put[0];
  free(put);
  return res;
}

double orc_crr_bermudan_put(double S, double K, double T, double r, double sigma, int dates, int steps_per_date) {
  const int steps = dates * steps_per_date;
  const double dt = T / steps;
  const double vdt = sigma * sqrt(dt);
  const double If = exp(r * dt), Df = exp(-r * dt);
  const double u = exp(vdt), d = exp(-vdt);
  const double pu = (If - d) / (u - d), pd = 1.0 - pu;
  const double puDf = pu * Df, pdDf = pd * Df;
  double* put = (double*)malloc(sizeof(double) * (steps + 1));
  for (int i = 0; i <= steps; ++i) {
    const double v = K - S * exp(vdt * (2.0 * i - steps));
    put[i] = v > 0.0 ? v : 0.0;
  }
  for (int i = steps; i > 0; --i) {
    const int layer = i - 1; /* time layer*dt of the nodes being computed */
    const int can_exercise = (layer % steps_per_date) == 0;
    for (int j = 0; j <= i - 1; ++j) {
      const double cont = puDf * put[j + 1] + pdDf * put[j];
      if (can_exercise) {
        const double ex = K - S * exp(vdt * (2.0 * j - layer));
        put[j] = cont > ex ? cont : ex;
      } else {
        put[j] = cont;
      }
    }
  }
  const double res = put[0];
  free(put);
  return res;
}

double orc_geometric_asian_call(double S, double X, double T, double r, double sigma, int N) {
  const double mu = log(S) + (r - 0.5 * sigma * sigma) * T * (N + 1.0) / (2.0 * N);
  const double var = sigma * sigma * T * (N + 1.0) * (2.0 * N + 1.0) / (6.0 * N * (double)N);
  const double sg = sqrt(var);
  const double d2 = (mu - log(X)) / sg;
  const double d1 = d2 + sg;
  return exp(-r * T) * (exp(mu + 0.5 * var) * ncdf(d1) - X * ncdf(d2));
}

/* ------------------------------------------------------------------------ */
/* GBM European, buffer mode: MonteCarloInsideBlockingDP.cpp:237-270.  The
 * samples already hold s = MuByT + VBySqrtT*z in the log2 domain (float);
 * payoff sums are sequential FP64 over paths. */
void orc_gbm_buffer(int n_opt, int n_paths, double r, double sigma, const double* S, const double* X,
                    const double* T, const float* samples, double* res, double* conf) {
  (void)sigma;
  const double RLog2E = -r * ORC_LOG2E;
  const double N = (double)n_paths;
  const double STDDEV_DENOM = 1.0 / (N * (N - 1.0));
  const double CONFIDENCE_DENOM = 1.96 / sqrt(N);
#pragma omp parallel for schedule(dynamic)
  for (int opt = 0; opt < n_opt; ++opt) {
    const double Y = S[opt], Z = X[opt];
    const float* s = samples + (size_t)opt * n_paths;
    double v0 = 0.0, v1 = 0.0;
    for (int i = 0; i < n_paths; ++i) {
      const double callValue = Y * exp2((double)s[i]) - Z;
      if (callValue > 0.0) {
        v0 += callValue;
        v1 += callValue * callValue;
      }
    }
    const double exprt = exp2(RLog2E * T[opt]);
    res[opt] = exprt * v0 / N;
    const double stdDev = sqrt((N * v1 - v0 * v0) * STDDEV_DENOM);
    conf[opt] = exprt * stdDev * CONFIDENCE_DENOM;
  }
}

/* Control-variate estimator (an engine option with no counterpart in the
 * samples): y = (S_T - K)^+, u = S_T - S e^{rT} (mean 0), price =
 * e^{-rT} (mean(y) - beta mean(u)), beta = cov(y, u) / var(u), confidence
 * 1.96 e^{-rT} sqrt((var(y) - beta cov(y, u)) / N).  Pinned by the small-K
 * limit (every path in the money: y = u + S e^{rT} - K exactly, so the price
 * is S - K e^{-rT} with zero variance) and by Black-Scholes. */
static void cv_price(double sy, double syy, double su, double suu, double suy, double N, double exprt,
                     double* res, double* conf) {
  const double d = 1.0 / (N * (N - 1.0));
  const double vy = (N * syy - sy * sy) * d;
  const double vu = (N * suu - su * su) * d;
  const double cov = (N * suy - su * sy) * d;
  const double beta = vu > 0.0 ? cov / vu : 0.0;
  const double v = vy - beta * cov;
  *res = exprt * (sy - beta * su) / N;
  *conf = exprt * sqrt(v > 0.0 ? v : 0.0) * (1.96 / sqrt(N));
}

void orc_gbm_buffer_cv(int n_opt, int n_paths, double r, const double* S, const double* X, const double* T,
                       const float* samples, double* res, double* conf) {
  const double N = (double)n_paths;
#pragma omp parallel for schedule(dynamic)
  for (int opt = 0; opt < n_opt; ++opt) {
    const double Y = S[opt], Z = X[opt], F = S[opt] * exp(r * T[opt]);
    const float* s = samples + (size_t)opt * n_paths;
    double sy = 0, syy = 0, su = 0, suu = 0, suy = 0;
    for (int i = 0; i < n_paths; ++i) {
      const double st = Y * exp2((double)s[i]);
      const double c = st - Z, u = st - F;
      if (c > 0.0) {
        sy += c;
        syy += c * c;
        suy += u * c;
      }
      su += u;
      suu += u * u;
    }
    cv_price(sy, syy, su, suu, suy, N, exp2(-r * ORC_LOG2E * T[opt]), &res[opt], &conf[opt]);
  }
}

void orc_gbm_philox_cv(int n_opt, int n_paths, double r, double sigma, const double* S, const double* X,
                       const double* T, uint64_t seed, uint32_t stream, double* res, double* conf) {
  const double N = (double)n_paths;
#pragma omp parallel for schedule(dynamic)
  for (int opt = 0; opt < n_opt; ++opt) {
    const double Tt = T[opt];
    const double mu = ORC_LOG2E * (r - 0.5 * sigma * sigma) * Tt;
    const double vol = ORC_LOG2E * sigma * sqrt(Tt);
    const double Y = S[opt], Z = X[opt], F = S[opt] * exp(r * Tt);
    double sy = 0, syy = 0, su = 0, suu = 0, suy = 0;
    for (int q = 0; q < (n_paths + 3) / 4; ++q) {
      double z[4];
      orc_normals4(seed, stream, 1u, (uint32_t)q, 0u, (uint32_t)opt, z);
      for (int j = 0; j < 4 && 4 * q + j < n_paths; ++j) {
        const double st = Y * exp2(mu + vol * z[j]);
        const double c = st - Z, u = st - F;
        if (c > 0.0) {
          sy += c;
          syy += c * c;
          suy += u * c;
        }
        su += u;
        suu += u * u;
      }
    }
    cv_price(sy, syy, su, suu, suy, N, exp2(-r * ORC_LOG2E * Tt), &res[opt], &conf[opt]);
  }
}

static void mean_se(double s, double s2, double N, double f, double* mean, double* se) {
  *mean = f * s / N;
  double var = (N * s2 - s * s) / (N * (N - 1.0));
  if (var < 0.0) var = 0.0;
  *se = fabs(f) * sqrt(var) / sqrt(N);
}

/* GBM European, Philox mode: same estimator; normals by the engine's Philox
 * contract (GBM: counter (path/4, 0, option), normal path%4).  Pathwise
 * Greeks: delta = e^{-rT} 1{S_T>K} S_T/S0, vega = e^{-rT} 1{S_T>K} S_T
 * (sqrt(T) z - sigma T), rho = e^{-rT} 1{S_T>K} K T. */
void orc_gbm_philox(int n_opt, int n_paths, double r, double sigma, const double* S, const double* X,
                    const double* T, uint64_t seed, uint32_t stream, double* res, double* conf,
                    double* greeks6) {
  const double N = (double)n_paths;
#pragma omp parallel for schedule(dynamic)
  for (int opt = 0; opt < n_opt; ++opt) {
    const double Tt = T[opt], sqrtT = sqrt(Tt);
    const double mu = ORC_LOG2E * (r - 0.5 * sigma * sigma) * Tt;
    const double vol = ORC_LOG2E * sigma * sqrtT;
    const double Y = S[opt], Z = X[opt];
    double v0 = 0, v1 = 0, dl = 0, dl2 = 0, vg = 0, vg2 = 0, itm = 0;
    for (int q = 0; q < (n_paths + 3) / 4; ++q) {
      double z[4];
      orc_normals4(seed, stream, 1u, (uint32_t)q, 0u, (uint32_t)opt, z);
      for (int j = 0; j < 4 && 4 * q + j < n_paths; ++j) {
        const double st = Y * exp2(mu + vol * z[j]);
        const double c = st - Z;
        if (c > 0.0) {
          const double g = st * (sqrtT * z[j] - sigma * Tt);
          v0 += c;
          v1 += c * c;
          dl += st;
          dl2 += st * st;
          vg += g;
          vg2 += g * g;
          itm += 1.0;
        }
      }
    }
    const double exprt = exp2(-r * ORC_LOG2E * Tt);
    double se;
    mean_se(v0, v1, N, exprt, &res[opt], &se);
    conf[opt] = 1.96 * se;
    if (greeks6) {
      double* g = greeks6 + 6 * (size_t)opt;
      mean_se(dl, dl2, N, exprt / Y, &g[0], &g[3]);
      mean_se(vg, vg2, N, exprt, &g[1], &g[4]);
      mean_se(itm, itm, N, exprt * Z * Tt, &g[2], &g[5]);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Heston Asian, buffer mode: MonteCarloTimeStep, MCAsianOptions.cpp:112-201
 * for one pass.  Per 64-path chunk the sample correlates the two streams
 * (:153-158), evolves the variance for all steps (:159 -> :65-74) and then
 * the spot with the start-of-step variance (:163-174 -> :51-62), averages
 * steps 1..N (:171,181), and accumulates the chunk mean (:184-192).
 * Buffers: z[((opt*nchunk + chunk)*n_steps + step)*64 + lane]. */
void orc_heston_buffer(int n_opt, int n_paths, int n_steps, double r, double rho, const double* S,
                       const double* X, const double* T, const double* v0, const double* kappa,
                       const double* theta, const double* xi, const double* z1, const double* z2,
                       uint32_t quirks, double* res, double* conf) {
  const int L = 64;
  const int nchunk = n_paths / L;
  const double N = (double)n_paths;
  const double STDDEV_DENOM = 1.0 / (N * (N - 1.0));
  const double CONFIDENCE_DENOM = 1.0 / sqrt(N);
#pragma omp parallel for schedule(dynamic)
  for (int opt = 0; opt < n_opt; ++opt) {
    const double dt = T[opt] / n_steps;
    const double sqrtRhoSq = sqrt(1.0 - rho * rho);
    double val = 0.0, val2 = 0.0, s0 = 0.0, s1 = 0.0;
    double* vol = (double*)malloc(sizeof(double) * (size_t)(n_steps + 1) * L);
    for (int ch = 0; ch < nchunk; ++ch) {
      const double* a = z1 + ((size_t)opt * nchunk + ch) * (size_t)n_steps * L;
      const double* b = z2 + ((size_t)opt * nchunk + ch) * (size_t)n_steps * L;
      for (int iv = 0; iv < L; ++iv) vol[iv] = v0[opt];
      for (int st = 0; st < n_steps; ++st) {
        for (int iv = 0; iv < L; ++iv) {
          const double zc = rho * a[st * L + iv] + sqrtRhoSq * b[st * L + iv];
          const double prev = vol[st * L + iv];
          const double vmax = prev > 0.0 ? prev : 0.0;
          vol[(st + 1) * L + iv] = prev + kappa[opt] * dt * (theta[opt] - vmax) + xi[opt] * sqrt(vmax * dt) * zc;
        }
      }
      double tmpVal = 0.0;
      for (int iv = 0; iv < L; ++iv) {
        double sim = S[opt], avg = 0.0;
        for (int st = 0; st < n_steps; ++st) {
          const double vmax = vol[st * L + iv] > 0.0 ? vol[st * L + iv] : 0.0;
          const double sv = (quirks & ORC_Q_VAR_AS_VOL) ? vmax : sqrt(vmax);
          const double MuByT = (r - 0.5 * sv * sv) * dt;
          const double VBySqrtT = sv * sqrt(dt);
          sim = sim * exp(MuByT + VBySqrtT * a[st * L + iv]);
          avg += sim;
        }
        double cv = avg / n_steps - X[opt];
        cv = cv > 0.0 ? cv : 0.0;
        tmpVal += cv;
        s0 += cv;
        s1 += cv * cv;
      }
      val += tmpVal / L;
      val2 += val * val;
    }
    free(vol);
    const double exprt = exp(-r * T[opt]);
    if (quirks & ORC_Q_RUNNING_VAL2) {
      res[opt] = exprt * val / N * L;
      const double stdDev = sqrt((N * val2 - val * val) * STDDEV_DENOM);
      conf[opt] = exprt * 1.96 * stdDev * CONFIDENCE_DENOM;
    } else {
      double se;
      mean_se(s0, s1, N, exprt, &res[opt], &se);
      conf[opt] = 1.96 * se;
    }
  }
}

/* Heston Asian, Philox mode, with pathwise Greeks by forward (tangent)
 * propagation -- an independent formulation of the derivatives the GPU
 * computes by a reverse sweep.  Normals: counter (path, step/2, option);
 * step 2j uses normals (0, 1), step 2j+1 normals (2, 3). */
void orc_heston_philox(int n_opt, int n_paths, int n_steps, double r, double rho, const double* S,
                       const double* X, const double* T, const double* v0, const double* kappa,
                       const double* theta, const double* xi, uint64_t seed, uint32_t stream,
                       uint32_t quirks, double* res, double* conf, double* greeks6) {
  const double N = (double)n_paths;
  const double srho = sqrt(1.0 - rho * rho);
#pragma omp parallel for schedule(dynamic)
  for (int opt = 0; opt < n_opt; ++opt) {
    const double dt = T[opt] / n_steps, sdt = sqrt(dt);
    const double disc = exp(-r * T[opt]);
    const double sig0 = sqrt(v0[opt]);
    double s0 = 0, s1 = 0, g[3] = {0, 0, 0}, g2[3] = {0, 0, 0};
    for (int p = 0; p < n_paths; ++p) {
      double Sx = S[opt], v = v0[opt], avg = 0.0;
      /* tangents: [0] d/dS0, [1] d/dsigma0 (v0 = sigma0^2), [2] d/dr */
      double dS[3] = {1.0, 0.0, 0.0}, dv[3] = {0.0, 2.0 * sig0, 0.0}, dA[3] = {0, 0, 0};
      double z[4] = {0, 0, 0, 0};
      for (int k = 0; k < n_steps; ++k) {
        if ((k & 1) == 0) orc_normals4(seed, stream, 2u, (uint32_t)p, (uint32_t)(k >> 1), (uint32_t)opt, z);
        const double za = (k & 1) ? z[2] : z[0];
        const double zb = (k & 1) ? z[3] : z[1];
        const double zc = rho * za + srho * zb;
        const double vp = v > 0.0 ? v : 0.0;
        const double sq = sqrt(vp);
        const double sv = (quirks & ORC_Q_VAR_AS_VOL) ? vp : sq;
        const double e = r * dt - 0.5 * sv * sv * dt + sv * sdt * za;
        const double growth = exp(e);
        const double Snew = Sx * growth;
        /* de/dv and dv'/dv on the v > 0 branch (correct convention). */
        const double dedv = vp > 0.0 ? (-0.5 * dt + sdt * za * 0.5 / sq) : 0.0;
        const double dvdv = vp > 0.0 ? (1.0 - kappa[opt] * dt + xi[opt] * sdt * zc * 0.5 / sq) : 1.0;
        for (int t = 0; t < 3; ++t) {
          const double de = dedv * dv[t] + (t == 2 ? dt : 0.0);
          dS[t] = dS[t] * growth + Snew * de;
          dv[t] = dv[t] * dvdv;
          dA[t] += dS[t];
        }
        v = v + kappa[opt] * dt * (theta[opt] - vp) + xi[opt] * sq * sdt * zc;
        Sx = Snew;
        avg += Sx;
      }
      const double x = avg / n_steps - X[opt];
      if (x > 0.0) {
        s0 += x;
        s1 += x * x;
        for (int t = 0; t < 3; ++t) {
          double d = disc * dA[t] / n_steps;
          if (t == 2) d += -T[opt] * disc * x;
          g[t] += d;
          g2[t] += d * d;
        }
      }
    }
    double se;
    mean_se(s0, s1, N, disc, &res[opt], &se);
    conf[opt] = 1.96 * se;
    if (greeks6)
      for (int t = 0; t < 3; ++t) mean_se(g[t], g2[t], N, 1.0, &greeks6[6 * opt + t], &greeks6[6 * opt + 3 + t]);
  }
}

/* ------------------------------------------------------------------------ */
/* LIBOR market model, liborSwaptionGreeks.cpp.  NN = 80 rates, NMAT = 40
 * steps, delta = 0.25 (params.h:27-34); portfolio of params.h:38-40 with the
 * swap rates rounded to float as the sample initialises them. */
#define ORC_NN 80
#define ORC_NMAT 40
static const int orc_mat[15] = {4, 4, 4, 8, 8, 8, 20, 20, 20, 28, 28, 28, 40, 40, 40};
static const float orc_rate[15] = {.045f, .05f, .055f, .045f, .05f, .055f, .045f, .05f,
                                   .055f, .045f, .05f, .055f, .045f, .05f, .055f};

double orc_libor_path(const double* z, const double* L0, double lambda, uint32_t quirks, double* Lb) {
  const double delta = 0.25, sqrt_delta = 0.5;
  double L[ORC_NN], L2[(ORC_NMAT + 1) * ORC_NN];
  for (int i = 0; i < ORC_NN; ++i) L[i] = L0[i];
  memset(L2, 0, sizeof(L2));
  /* path_calc_b1, :34-82 */
  for (int i = 0; i < ORC_NN; ++i) L2[i] = L[i];
  for (int n = 0; n < ORC_NMAT; ++n) {
    const double sqez = sqrt_delta * z[n];
    double v = 0.0;
    for (int i = n + 1; i < ORC_NN; ++i) {
      const double lam = lambda, con1 = delta * lam;
      v += con1 * L[i] / (1.0 + delta * L[i]);
      const double vrat = exp(con1 * v + lam * (sqez - 0.5 * con1));
      L[i] = L[i] * vrat;
      L2[i + (n + 1) * ORC_NN] = L[i];
    }
  }
  /* portfolio_b, :120-211; with ORC_Q_LIBOR_INPLACE the adjoint overwrites L
   * exactly as the sample does (:178, read back at :181). */
  double B[ORC_NMAT], S[ORC_NMAT], B_b[ORC_NMAT], S_b[ORC_NMAT];
  double b = 1.0, s = 0.0, v = 0.0;
  double Lw[ORC_NN];
  memcpy(Lw, L, sizeof(L));
  for (int m = 0; m < ORC_NN - ORC_NMAT; ++m) {
    const int n = m + ORC_NMAT;
    b = b / (1.0 + delta * Lw[n]);
    s = s + delta * b;
    B[m] = b;
    S[m] = s;
  }
  for (int i = 0; i < ORC_NMAT; ++i) B_b[i] = S_b[i] = 0.0;
  for (int n = 0; n < 15; ++n) {
    const int m = orc_mat[n] - 1;
    const double K = (double)orc_rate[n];
    const double swapval = B[m] + K * S[m] - 1.0;
    if (swapval < 0.0) {
      v += -100.0 * swapval;
      S_b[m] += -100.0 * K;
      B_b[m] += -100.0;
    }
  }
  double Lbar[ORC_NN];
  for (int m = ORC_NN - ORC_NMAT - 1; m >= 0; --m) {
    const int n = m + ORC_NMAT;
    B_b[m] += delta * S_b[m];
    Lbar[n] = -B_b[m] * B[m] * delta / (1.0 + delta * Lw[n]);
    if (quirks & ORC_Q_LIBOR_INPLACE) Lw[n] = Lbar[n];
    if (m > 0) {
      S_b[m - 1] += S_b[m];
      B_b[m - 1] += B_b[m] / (1.0 + delta * Lw[n]);
    }
  }
  b = 1.0;
  for (int n = 0; n < ORC_NMAT; ++n) b = b / (1.0 + delta * Lw[n]);
  v = b * v;
  for (int n = 0; n < ORC_NMAT; ++n) Lbar[n] = -v * delta / (1.0 + delta * Lw[n]);
  for (int n = ORC_NMAT; n < ORC_NN; ++n) Lbar[n] = b * Lbar[n];
  /* path_calc_b2, :86-117 */
  for (int n = ORC_NMAT - 1; n >= 0; --n) {
    double v1 = 0.0;
    for (int i = ORC_NN - 1; i > n; --i) {
      const double lam = lambda;
      const double l2n = L2[i + n * ORC_NN], l2n1 = L2[i + (n + 1) * ORC_NN];
      v1 += lam * l2n1 * Lbar[i];
      const double faci = delta / (1.0 + delta * l2n);
      Lbar[i] = Lbar[i] * l2n1 / l2n + v1 * lam * faci * faci;
    }
  }
  memcpy(Lb, Lbar, sizeof(Lbar));
  return v;
}

void orc_libor(int n_paths, double L0, double lambda, const double* z, uint64_t seed, uint32_t stream,
               uint32_t quirks, double* value, double* deltas) {
  double s0 = 0, s1 = 0, d0[ORC_NN], d1[ORC_NN], curve[ORC_NN];
  for (int i = 0; i < ORC_NN; ++i) {
    d0[i] = d1[i] = 0.0;
    curve[i] = L0;
  }
#pragma omp parallel
  {
    double t0 = 0, t1 = 0, e0[ORC_NN], e1[ORC_NN], zp[ORC_NMAT], Lb[ORC_NN];
    for (int i = 0; i < ORC_NN; ++i) e0[i] = e1[i] = 0.0;
#pragma omp for schedule(static)
    for (int p = 0; p < n_paths; ++p) {
      if (z) {
        memcpy(zp, z + (size_t)p * ORC_NMAT, sizeof(zp));
      } else {
        for (int q = 0; q < ORC_NMAT / 4; ++q) orc_normals4(seed, stream, 3u, (uint32_t)p, (uint32_t)q, 0u, zp + 4 * q);
      }
      const double v = orc_libor_path(zp, curve, lambda, quirks, Lb);
      t0 += v;
      t1 += v * v;
      for (int i = 0; i < ORC_NN; ++i) {
        e0[i] += Lb[i];
        e1[i] += Lb[i] * Lb[i];
      }
    }
#pragma omp critical
    {
      s0 += t0;
      s1 += t1;
      for (int i = 0; i < ORC_NN; ++i) {
        d0[i] += e0[i];
        d1[i] += e1[i];
      }
    }
  }
  const double N = (double)n_paths;
  mean_se(s0, s1, N, 1.0, &value[0], &value[1]);
  for (int i = 0; i < ORC_NN; ++i) mean_se(d0[i], d1[i], N, 1.0, &deltas[i], &deltas[ORC_NN + i]);
}

/* ------------------------------------------------------------------------ */
/* Longstaff-Schwartz, reference algorithm: MonteCarloAmericanOptions/main.cpp
 * :46-60 (constants), utilities.cpp:42-88 (perform_sims, log2 domain),
 * main.cpp:131-279 (backward sweep) and utilities.cpp:90-123 (invert3_arr).
 * z[(s-1)*n_paths + p] drives step s = 1..n_steps of path p. */
static double det3(double a, double b, double c, double d, double e, double f, double g, double h, double i) {
  return a * e * i + d * h * c + g * b * f - g * e * c - d * b * i - a * h * f;
}

static void invert3(double a[3][3], double x[3][3]) {
  const double det = det3(x[0][0], x[0][1], x[0][2], x[1][0], x[1][1], x[1][2], x[2][0], x[2][1], x[2][2]);
  a[0][0] = (x[1][1] * x[2][2] - x[2][1] * x[1][2]) / det;
  a[0][1] = (x[0][2] * x[2][1] - x[0][1] * x[2][2]) / det;
  a[0][2] = (x[0][1] * x[1][2] - x[0][2] * x[1][1]) / det;
  a[1][0] = (x[1][2] * x[2][0] - x[1][0] * x[2][2]) / det;
  a[1][1] = (x[0][0] * x[2][2] - x[0][2] * x[2][0]) / det;
  a[1][2] = (x[0][2] * x[1][0] - x[0][0] * x[1][2]) / det;
  a[2][0] = (x[1][0] * x[2][1] - x[1][1] * x[2][0]) / det;
  a[2][1] = (x[0][1] * x[2][0] - x[0][0] * x[2][1]) / det;
  a[2][2] = (x[0][0] * x[1][1] - x[0][1] * x[1][0]) / det;
}

void orc_lsmc_ref(int n_paths, int n_steps, double spot, double strike, double vol, double maturity,
                  const double* z, double* american, double* european, double* betas) {
  const double dt = maturity / n_steps;
  const double volSqrtdt = vol * sqrt(dt);
  const double fwdFactor = vol * vol * dt * -0.5;
  const double vsqrtdt_log2e = volSqrtdt * ORC_LOG2E;
  const double fwdFactor_log2e = fwdFactor * ORC_LOG2E;
  const double avgPathFactor = 1.0 / n_paths;
  const size_t N = (size_t)n_paths;
  double* a = (double*)malloc(sizeof(double) * (size_t)(n_steps + 1) * N);
  double* cf = (double*)calloc(N, sizeof(double));
  for (size_t p = 0; p < N; ++p) a[p] = spot;
  for (int s = 1; s <= n_steps; ++s)
    for (size_t p = 0; p < N; ++p)
      a[s * N + p] = a[(s - 1) * N + p] * exp2(fwdFactor_log2e + vsqrtdt_log2e * z[(size_t)(s - 1) * N + p]);
  double b0 = 0.0, b1 = 0.0, b2 = 0.0;
  for (int s = n_steps - 1; s > 0; --s) {
    double sum0 = 0, sum1 = 0, sum2 = 0, sum3 = 0, sum4 = 0, x0 = 0, x1 = 0, x2 = 0;
    const double* ap = a + (size_t)s * N;
    for (size_t p = 0; p < N; ++p) {
      const double x = ap[p];
      if (s < n_steps - 1) {
        if (x <= strike) {
          const double reg = b0 + b1 * x + b2 * (x * x);
          const double payoff = strike - x;
          if (payoff > reg) cf[p] = payoff;
        }
      } else {
        cf[p] = (strike - x > 0.0 ? strike - x : 0.0) * avgPathFactor;
      }
      if (x <= strike) {
        const double x2v = x * x;
        sum0 += 1;
        sum1 += x;
        sum2 += x2v;
        sum3 += x2v * x;
        sum4 += x2v * x2v;
        x0 += cf[p];
        x1 += cf[p] * x;
        x2 += cf[p] * x2v;
      }
    }
    const double ridge = 0.01;
    sum0 += ridge;
    sum1 += ridge;
    sum2 += ridge;
    sum3 += ridge;
    sum4 += ridge;
    double m[3][3] = {{sum0, sum1, sum2}, {sum1, sum2, sum3}, {sum2, sum3, sum4}}, inv[3][3];
    invert3(inv, m);
    b0 = x0 * inv[0][0] + x1 * inv[1][0] + x2 * inv[2][0];
    b1 = x0 * inv[0][1] + x1 * inv[1][1] + x2 * inv[2][1];
    b2 = x0 * inv[0][2] + x1 * inv[1][2] + x2 * inv[2][2];
    if (betas) {
      betas[3 * s] = b0;
      betas[3 * s + 1] = b1;
      betas[3 * s + 2] = b2;
    }
  }
  double am = 0.0, eu = 0.0;
  for (size_t p = 0; p < N; ++p) {
    am += cf[p];
    const double e = strike - a[(size_t)n_steps * N + p];
    eu += e > 0.0 ? e : 0.0;
  }
  *american = am / n_paths;
  *european = eu / n_paths;
  free(a);
  free(cf);
}

int orc_chol4_solve(const double* G, const double* rhs, double* beta) {
  double L[4][4] = {{0}};
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j <= i; ++j) {
      double s = G[4 * i + j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (!(s > 1e-13 * G[4 * i + i])) return 0;
        L[i][i] = sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  }
  double y[4];
  for (int i = 0; i < 4; ++i) {
    double s = rhs[i];
    for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
    y[i] = s / L[i][i];
  }
  for (int i = 3; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < 4; ++k) s -= L[k][i] * beta[k];
    beta[i] = s / L[i][i];
  }
  return 1;
}

/* Correct LSMC for one option.  Normals: counter (path, j/4, option) of the
 * LSMC tag, normal j%4 drives the step from S_j to S_{j+1}. */
static void lsmc_one(int o, double S0, double K, double sigma, double r, double T, int n, int P, uint64_t seed,
                     uint32_t stream, double* price2, double* european) {
  /* Paths by the Brownian bridge, backwards from T (the engine's layout,
   * documented in include/fsmc.h): Z_i of path p = component p & 3 of
   * counter (p / 4, i, option), W_N = sqrt(T) Z_{N-1},
   * W_k = k/(k+1) W_{k+1} + sqrt(dt k/(k+1)) Z_{k-1},
   * S_k = S0 exp((r - sigma^2/2) dt k + sigma W_k), S_N with (r - sigma^2/2) T. */
  const double dt = T / n;
  const double mu = (r - 0.5 * sigma * sigma) * dt;
  double* S = (double*)malloc(sizeof(double) * (size_t)n * P); /* row j = S_{j+1} */
  double* W = (double*)malloc(sizeof(double) * (size_t)n);     /* W[k-1] = W_k of one path */
  double* cf = (double*)malloc(sizeof(double) * (size_t)P);
  const double discT = exp(-r * T);
  double eu = 0.0;
  for (int p = 0; p < P; ++p) {
    double zq[4];
    for (int k = n; k >= 1; --k) {
      const int i = k - 1;
      orc_normals4(seed, stream, 4u, (uint32_t)(p >> 2), (uint32_t)i, (uint32_t)o, zq);
      const double z = zq[p & 3];
      W[k - 1] = k == n ? sqrt(T) * z : ((double)k / (k + 1)) * W[k] + sqrt(dt * k / (k + 1)) * z;
    }
    for (int k = 1; k < n; ++k) S[(size_t)(k - 1) * P + p] = S0 * exp(mu * k + sigma * W[k - 1]);
    const double x = S0 * exp((r - 0.5 * sigma * sigma) * T + sigma * W[n - 1]);
    S[(size_t)(n - 1) * P + p] = x;
    const double pay = K - x > 0.0 ? K - x : 0.0;
    cf[p] = discT * pay;
    eu += cf[p];
  }
  free(W);
  for (int k = n - 1; k >= 1; --k) {
    const double tk = k * dt, disc = exp(-r * tk);
    const double inv_sw = 1.0 / (sigma * sqrt(tk));
    const double* row = S + (size_t)(k - 1) * P;
    double mo[7] = {0}, c[4] = {0};
    for (int p = 0; p < P; ++p) {
      if (K - row[p] > 0.0) {
        const double x = 1.0 + (row[p] / K - 1.0) * inv_sw;
        double xp = 1.0;
        for (int q = 0; q < 7; ++q) {
          mo[q] += xp;
          if (q < 4) c[q] += xp * cf[p];
          xp *= x;
        }
      }
    }
    double G[16], beta[4];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) G[4 * i + j] = mo[i + j];
    if (mo[0] < 4.0 || !orc_chol4_solve(G, c, beta)) continue; /* no exercise at this date */
    for (int p = 0; p < P; ++p) {
      if (K - row[p] > 0.0) {
        const double x = 1.0 + (row[p] / K - 1.0) * inv_sw;
        const double cont = beta[0] + x * (beta[1] + x * (beta[2] + x * beta[3]));
        const double ex = (K - row[p]) * disc;
        if (ex > cont) cf[p] = ex;
      }
    }
  }
  double s0 = 0.0, s1 = 0.0;
  for (int p = 0; p < P; ++p) {
    s0 += cf[p];
    s1 += cf[p] * cf[p];
  }
  double mean, se;
  mean_se(s0, s1, (double)P, 1.0, &mean, &se);
  const double now = K - S0 > 0.0 ? K - S0 : 0.0;
  price2[0] = mean > now ? mean : now;
  price2[1] = se;
  if (european) *european = eu / P;
  free(S);
  free(cf);
}

void orc_lsmc(int n_opt, const double* S0, const double* K, const double* sigma, const double* r,
              const double* T, const int* steps, const int* paths, uint64_t seed, uint32_t stream,
              double* price2, double* european) {
#pragma omp parallel for schedule(dynamic)
  for (int o = 0; o < n_opt; ++o)
    lsmc_one(o, S0[o], K[o], sigma[o], r[o], T[o], steps[o], paths[o], seed, stream, price2 + 2 * o,
             european ? european + o : NULL);
}

/* ==========================================================================
 * MKL-compatible MRG32k3a + inverse-CDF normals (mklRng.cpp:22-32, 44-139;
 * mklRng.h:9).  Recurrences of L'Ecuyer, "Good parameters and implementations
 * for combined multiple recursive random number generators", Oper. Res. 47
 * (1999):  x_n = (1403580 x_{n-2} - 810728 x_{n-3}) mod m1,
 *          y_n = (527612 y_{n-1} - 1370589 y_{n-3}) mod m2.
 * Skip-ahead by matrix powers of the one-step transitions.
 * ========================================================================== */
#define ORC_M1 4294967087ULL
#define ORC_M2 4294944443ULL

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) {
  return (uint64_t)(((unsigned __int128)a * b) % m);
}

/* C = A B (3x3, mod m) */
static void mat3_mul(const uint64_t A[9], const uint64_t B[9], uint64_t m, uint64_t C[9]) {
  uint64_t T[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      uint64_t acc = 0;
      for (int k = 0; k < 3; ++k) acc = (acc + mulmod(A[3 * i + k], B[3 * k + j], m)) % m;
      T[3 * i + j] = acc;
    }
  memcpy(C, T, sizeof(T));
}

static void mat3_vec(const uint64_t A[9], const uint64_t v[3], uint64_t m, uint64_t out[3]) {
  uint64_t t[3];
  for (int i = 0; i < 3; ++i) {
    uint64_t acc = 0;
    for (int k = 0; k < 3; ++k) acc = (acc + mulmod(A[3 * i + k], v[k], m)) % m;
    t[i] = acc;
  }
  memcpy(out, t, sizeof(t));
}

/* state (s[0], s[1], s[2]) = (x_{n-3}, x_{n-2}, x_{n-1}) advanced by k steps */
static void mrg_skip(uint64_t x[3], uint64_t y[3], uint64_t k) {
  uint64_t A1[9] = {0, 1, 0, 0, 0, 1, ORC_M1 - 810728ULL, 1403580ULL, 0};
  uint64_t A2[9] = {0, 1, 0, 0, 0, 1, ORC_M2 - 1370589ULL, 0, 527612ULL};
  while (k) {
    if (k & 1) {
      mat3_vec(A1, x, ORC_M1, x);
      mat3_vec(A2, y, ORC_M2, y);
    }
    mat3_mul(A1, A1, ORC_M1, A1);
    mat3_mul(A2, A2, ORC_M2, A2);
    k >>= 1;
  }
}

void orc_mrg32k3a_uniform(uint32_t seed, uint64_t offset, int64_t n, double* u) {
  uint64_t x[3] = {(uint64_t)seed % ORC_M1, 1, 1}, y[3] = {1, 1, 1};
  mrg_skip(x, y, offset);
  const double inv_m1 = 1.0 / 4294967087.0;
  for (int64_t i = 0; i < n; ++i) {
    /* -a x mod m = a (m - x): every product stays below 2^53 */
    const uint64_t xn = (1403580ULL * x[1] + 810728ULL * (ORC_M1 - x[0])) % ORC_M1;
    const uint64_t yn = (527612ULL * y[2] + 1370589ULL * (ORC_M2 - y[0])) % ORC_M2;
    x[0] = x[1];
    x[1] = x[2];
    x[2] = xn;
    y[0] = y[1];
    y[1] = y[2];
    y[2] = yn;
    const uint64_t z = (xn + ORC_M1 - yn % ORC_M1) % ORC_M1;
    /* z = 0 (probability 2^-32) would give u = 0; it is taken half a step up */
    u[i] = z ? (double)z * inv_m1 : 0.5 * inv_m1;
  }
}

/* Wichura, "Algorithm AS 241: The percentage points of the normal
 * distribution", Appl. Statist. 37 (1988) 477-484, PPND16. */
double orc_icdf(double p) {
  const double q = p - 0.5;
  if (fabs(q) <= 0.425) {
    const double r = 0.180625 - q * q;
    const double num = ((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r +
                            6.7265770927008700853e+4) * r + 4.5921953931549871457e+4) * r +
                          1.3731693765509461125e+4) * r + 1.9715909503065514427e+3) * r +
                        1.3314166789178437745e+2) * r + 3.3871328727963666080e+0;
    const double den = ((((((5.2264952788528545610e+3 * r + 2.8729085735721942674e+4) * r +
                            3.9307895800092710610e+4) * r + 2.1213794301586595867e+4) * r +
                          5.3941960214247511077e+3) * r + 6.8718700749205790830e+2) * r +
                        4.2313330701600911252e+1) * r + 1.0;
    return q * num / den;
  }
  double r = q < 0.0 ? p : 1.0 - p;
  r = sqrt(-log(r));
  double x;
  if (r <= 5.0) {
    r -= 1.6;
    const double num = ((((((7.74545014278341407640e-4 * r + 2.27238449892691845833e-2) * r +
                            2.41780725177450611770e-1) * r + 1.27045825245236838258e+0) * r +
                          3.64784832476320460504e+0) * r + 5.76949722146069140550e+0) * r +
                        4.63033784615654529590e+0) * r + 1.42343711074968357734e+0;
    const double den = ((((((1.05075007164441684324e-9 * r + 5.47593808499534494600e-4) * r +
                            1.51986665636164571966e-2) * r + 1.48103976427480074590e-1) * r +
                          6.89767334985100004550e-1) * r + 1.67638483018380384940e+0) * r +
                        2.05319162663775882187e+0) * r + 1.0;
    x = num / den;
  } else {
    r -= 5.0;
    const double num = ((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r +
                            1.24266094738807843860e-3) * r + 2.65321895265761230930e-2) * r +
                          2.96560571828504891230e-1) * r + 1.78482653991729133580e+0) * r +
                        5.46378491116411436990e+0) * r + 6.65790464350110377720e+0;
    const double den = ((((((2.04426310338993978564e-15 * r + 1.42151175831644588870e-7) * r +
                            1.84631831751005468180e-5) * r + 7.86869131145613259100e-4) * r +
                          1.48753612908506148525e-2) * r + 1.36929880922735805310e-1) * r +
                        5.99832206555887937690e-1) * r + 1.0;
    x = num / den;
  }
  return q < 0.0 ? -x : x;
}

void orc_mrg32k3a_gaussian(uint32_t seed, uint64_t offset, int64_t n, double mean, double stddev, double* z) {
  orc_mrg32k3a_uniform(seed, offset, n, z);
  for (int64_t i = 0; i < n; ++i) z[i] = mean + stddev * orc_icdf(z[i]);
}

/* ---- MT19937 as MKL's VSL_BRNG_MT19937 (AsianOptionsPricing/mkl_rng.h:5-69,
 * MCAsianOptions.cpp:92-93): Matsumoto-Nishimura MT19937 seeded by
 * init_by_array({seed}); outputs offset .. offset+n-1 of the stream. */
#define ORC_MT_N 624
#define ORC_MT_M 397

static void orc_mt_seed(uint32_t seed, uint32_t* mt) {
  mt[0] = 19650218u;
  for (int i = 1; i < ORC_MT_N; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  int i = 1;
  for (int k = ORC_MT_N; k > 0; --k) {
    mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + seed;
    if (++i >= ORC_MT_N) {
      mt[0] = mt[ORC_MT_N - 1];
      i = 1;
    }
  }
  for (int k = ORC_MT_N - 1; k > 0; --k) {
    mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
    if (++i >= ORC_MT_N) {
      mt[0] = mt[ORC_MT_N - 1];
      i = 1;
    }
  }
  mt[0] = 0x80000000u;
}

static void orc_mt_twist(uint32_t* mt) {
  /* in place, word by word in index order: word i reads the already-updated
   * words below it, exactly as the sequential generator does */
  for (int i = 0; i < ORC_MT_N; ++i) {
    const uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % ORC_MT_N] & 0x7fffffffu);
    mt[i] = mt[(i + ORC_MT_M) % ORC_MT_N] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
  }
}

void orc_mt19937_bits(uint32_t seed, uint64_t offset, int64_t n, uint32_t* out) {
  uint32_t mt[ORC_MT_N];
  orc_mt_seed(seed, mt);
  int pos = ORC_MT_N;
  for (uint64_t k = 0; k < offset + (uint64_t)n; ++k) {
    if (pos == ORC_MT_N) {
      orc_mt_twist(mt);
      pos = 0;
    }
    uint32_t y = mt[pos++];
    if (k < offset) continue;
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    out[k - offset] = y;
  }
}

/* Box-Muller pairs of consecutive outputs (VSL_RNG_METHOD_GAUSSIAN_BOXMULLER2,
 * mkl_rng.h:57-62): u = bits 2^-32 (u1 = 0 taken as 2^-33),
 * (z0, z1) = sqrt(-2 ln u1) (sin 2 pi u2, cos 2 pi u2); offset and n even. */
void orc_mt19937_gaussian(uint32_t seed, uint64_t offset, int64_t n, double mean, double stddev, double* z) {
  uint32_t* bits = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(uint32_t));
  orc_mt19937_bits(seed, offset, n, bits);
  for (int64_t j = 0; j + 1 < n; j += 2) {
    const double u1 = bits[j] ? (double)bits[j] * 0x1.0p-32 : 0x1.0p-33;
    const double u2 = (double)bits[j + 1] * 0x1.0p-32;
    const double r = sqrt(-2.0 * log(u1));
    z[j] = mean + stddev * (r * sin(2.0 * ORC_PI * u2));
    z[j + 1] = mean + stddev * (r * cos(2.0 * ORC_PI * u2));
  }
  free(bits);
}

/* Threads of the OpenMP loops over options (bench.py's CPU baseline times
 * all host threads and one thread); returns the previous setting. */
int orc_set_threads(int n) {
#ifdef _OPENMP
  const int prev = omp_get_max_threads();
  if (n > 0) omp_set_num_threads(n);
  return prev;
#else
  (void)n;
  return 1;
#endif
}

/*
 * CPU oracle of the Monte Carlo engine -- TEST INFRASTRUCTURE ONLY.
 *
 * Used by tests/ as the checker of the GPU kernels and by bench.py as the
 * timed CPU baseline.  Plain FP64 C; each function cites the reference
 * file:line whose algorithm it restates.  Parity pins of this file (see the
 * tests named in each comment):
 *   philox        Random123 known-answer vectors        tests/test_oracle_pins.py
 *   box_muller    moment and identity checks           tests/test_oracle_pins.py
 *   bs/crr        textbook values, convergence, parity tests/test_oracle_pins.py
 *   gbm_*         Black-Scholes price/Greeks within 3 SE
 *   heston_*      exact martingale mean, N=1 reduction to Black-Scholes,
 *                 N=2 conditional-closed-form quadrature (rho, xi, kappa,
 *                 theta, truncation, both vol conventions)
 */
#include "fsmc_oracle.h"

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
  const double u1 = (double)((a >> 8) + 1u) * 0x1.0p-24;
  const double u2 = (double)(b >> 8) * 0x1.0p-24;
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

// Per-path models shared by the per-workload kernels and the mixed-batch
// kernels, so that a request priced inside a batch follows exactly the
// arithmetic (and Philox counters) of the standalone entry point.
#pragma once

#include "fsmc_device.hpp"

namespace fsmc {

constexpr double kPathLog2e = 1.4426950408889634074;

// ---------------------------------------------------------------------------
// Heston (MCAsianOptions.cpp:51-74 with sqrt(v+) as the volatility)
// ---------------------------------------------------------------------------
template <typename Real>
struct HestonConsts {
  Real c_rdt, c_hdt, c_sdt, kdt, xsdt, theta, rho, srho, xr, xs;
  Real dt_n, hdt_n, sdt_n, one_kdt;  // natural-log units, for the tangent sweep
  __device__ HestonConsts(double T, int n_steps, double r, double kappa, double xi, double th, double rh) {
    const double dt = T / n_steps;
    const double sdt = sqrt(dt);
    dt_n = (Real)dt;
    hdt_n = (Real)(0.5 * dt);
    sdt_n = (Real)sdt;
    one_kdt = (Real)(1.0 - kappa * dt);
    c_rdt = (Real)(r * dt * kPathLog2e);
    c_hdt = (Real)(0.5 * dt * kPathLog2e);
    c_sdt = (Real)(sdt * kPathLog2e);
    kdt = (Real)(kappa * dt);
    xsdt = (Real)(xi * sdt);
    theta = (Real)th;
    rho = (Real)rh;
    srho = (Real)sqrt(1.0 - rh * rh);
    xr = (Real)(xi * sdt * rh);                   // xi sqrt(dt) rho
    xs = (Real)(xi * sdt * sqrt(1.0 - rh * rh));  // xi sqrt(dt) sqrt(1 - rho^2)
  }
};

// One full-truncation Euler step: the spot moves with the variance at the
// start of the step, then the variance is advanced:
//   S *= 2^(log2e ((r - sv^2/2) dt + sv sqrt(dt) z1)),  sv = sqrt(v+) (v+ with
//   the sample's variance-as-volatility convention),
//   v += kappa dt (theta - v+) + xi sqrt(v+ dt) (rho z1 + sqrt(1-rho^2) z2).
// With sv = sqrt(v+), sv^2 is taken as v+ itself.  Contraction is off and the
// fused multiply-adds are explicit, so every kernel that inlines the step
// (forward, adjoint forward sweep, reverse-sweep recompute, batch) rounds
// exactly alike whatever the optimiser does around it.
template <typename Real, bool kVarAsVol>
__device__ __forceinline__ void heston_step(const HestonConsts<Real>& c, Real& S, Real& v, Real z1, Real z2) {
#pragma clang fp contract(off)
  const Real vp = positive_part(v);
  const Real sq = fsqrt(vp);
  const Real sv = kVarAsVol ? vp : sq;
  const Real sv2 = kVarAsVol ? vp * vp : vp;
  const Real e = ffma(c.c_sdt * sv, z1, ffma(-c.c_hdt, sv2, c.c_rdt));
  S = S * fexp2(e);
  v = ffma(sq, ffma(c.xr, z1, c.xs * z2), ffma(c.kdt, c.theta - vp, v));
}

// The FP32 Philox step with its normals in polar form, z1 = r cs, z2 = r sn
// (box_muller_polar gives r2 = r^2): both normals only ever appear
// multiplied by sqrt(v+), so q = sqrt(v+ r^2) takes the radius inside the
// step's one square root (a transcendental issues at half the FMA rate).
template <typename Real>
__device__ __forceinline__ void heston_step_polar(const HestonConsts<Real>& c, Real& S, Real& v, Real r2, Real cs,
                                                  Real sn) {
#pragma clang fp contract(off)
  const Real vp = positive_part(v);
  const Real q = fsqrt(vp * r2);
  const Real e = ffma(c.c_sdt * q, cs, ffma(-c.c_hdt, vp, c.c_rdt));
  S = S * fexp2(e);
  v = ffma(q, ffma(c.xr, cs, c.xs * sn), ffma(c.kdt, c.theta - vp, v));
}

// Two consecutive steps sharing one Philox call (normals n0..n3); the running
// sum of the monitored spots takes the pair's sum (one FP32 rounding) in one
// FP64 add.
template <typename Real, bool kVarAsVol>
__device__ __forceinline__ void heston_pair(const HestonConsts<Real>& c, Real& S, Real& v, PathSum& avg,
                                            Real n0, Real n1, Real n2, Real n3, bool second) {
#pragma clang fp contract(off)
  heston_step<Real, kVarAsVol>(c, S, v, n0, n1);
  if (second) {
    const Real s1 = S;
    heston_step<Real, kVarAsVol>(c, S, v, n2, n3);
    avg.add(s1 + S);
  } else {
    avg.add(S);
  }
}

// Two consecutive Philox-mode steps from counter ctr: FP32 with the
// square-root variance takes the polar form, everything else the normals.
// Every kernel that steps Philox-mode paths goes through here, so the
// forward, adjoint and batch kernels see the same bits.
template <typename Real, bool kVarAsVol>
__device__ __forceinline__ void heston_philox_pair(const HestonConsts<Real>& c, Real& S, Real& v, PathSum& avg,
                                                   U4 ctr, uint32_t k0, uint32_t k1, bool second) {
#pragma clang fp contract(off)
  if constexpr (sizeof(Real) == 4 && !kVarAsVol) {
    const U4 w = philox4x32_10(ctr, k0, k1);
    float q0, c0, s0, q1, c1, s1;  // q0, q1: the squared radii
    box_muller_polar(w.x, w.y, q0, c0, s0);
    box_muller_polar(w.z, w.w, q1, c1, s1);
    heston_step_polar(c, S, v, q0, c0, s0);
    if (second) {
      const Real sa = S;
      heston_step_polar(c, S, v, q1, c1, s1);
      avg.add(sa + S);
    } else {
      avg.add(S);
    }
  } else {
    Real n0, n1, n2, n3;
    normals4<Real>(ctr, k0, k1, n0, n1, n2, n3);
    heston_pair<Real, kVarAsVol>(c, S, v, avg, n0, n1, n2, n3, second);
  }
}

// One step of heston_step_polar (FP32, polar normals) or heston_step (FP64,
// normals z1 = cs, z2 = sn, r2 = 1) that also returns the step's derivatives
// de/dv and dv'/dv at its starting variance (0 and 1 where the variance is
// clamped), formed as heston_step_tangent forms them: the adjoint's segment
// recompute keeps these instead of the step's inputs.
template <typename Real>
__device__ __forceinline__ void heston_step_derivs(const HestonConsts<Real>& c, Real& S, Real& v, Real r2, Real cs,
                                                   Real sn, Real& dedv, Real& dvdv) {
#pragma clang fp contract(off)
  const Real vp = positive_part(v);
  Real q, u;  // q: the step's sqrt(vp) times the radius; u = radius / (2 sqrt(vp))
  if constexpr (sizeof(Real) == 4) {
    q = fsqrt(vp * r2);
    u = vp > Real(0) ? Real(0.5) * q * frcp(vp) : Real(0);
  } else {
    q = fsqrt(vp);
    u = vp > Real(0) ? Real(0.5) / q : Real(0);
  }
  const Real e = ffma(c.c_sdt * q, cs, ffma(-c.c_hdt, vp, c.c_rdt));
  const Real zz = ffma(c.xr, cs, c.xs * sn);
  S = S * fexp2(e);
  dedv = vp > Real(0) ? ffma(c.sdt_n * u, cs, -c.hdt_n) : Real(0);
  dvdv = vp > Real(0) ? ffma(zz, u, c.one_kdt) : Real(1);
  v = ffma(q, zz, ffma(c.kdt, c.theta - vp, v));
}

// Adjoint (reverse sweep) of one Heston-Asian path.  Pathwise derivatives of
//   P = e^{-rT} max(A - K, 0),  A = (1/N) sum_{k=1..N} S_k,
//   S_{k+1} = S_k exp(e_k),  e_k = r dt - vp_k dt / 2 + sqrt(vp_k dt) z1_k,
//   v_{k+1} = v_k + kappa dt (theta - vp_k) + xi sqrt(vp_k dt) zc_k,
// with vp = max(v, 0).  Going backwards with U_{k+1} = dP/de_k =
// (e^{-rT}/N) sum_{j>k} S_j:
//   vbar_k = vbar_{k+1} dv_{k+1}/dv_k + U_{k+1} de_k/dv_k,
//   rbar   = -T P + dt sum_k U_{k+1},   dP/dS0 = U_1 / S0,
// and vega = dP/dsqrt(v0) = 2 sqrt(v0) vbar_0.  The forward sweep stores
// (S, v) every kHestonSeg steps at ck[(2 seg) * stride] / ck[(2 seg + 1) *
// stride]; the reverse sweep recomputes one segment at a time into registers,
// regenerating its normals from the Philox counters (path, step/2, c2).
// Returns the undiscounted payoff; g = {dP/dS0, dP/dsqrt(v0), dP/dr}.
constexpr int kHestonSeg = 16;  // even: segments start on a Philox pair

template <typename Real>
__device__ __forceinline__ double heston_adjoint_path(const HestonConsts<Real>& c, double S0d, double Kd, double Td,
                                                      double v0d, double r, int N, uint32_t p, uint32_t c2,
                                                      uint32_t w, uint32_t k0, uint32_t k1, Real* ck, size_t stride,
                                                      double (&g)[3]) {
  const int nseg = (N + kHestonSeg - 1) / kHestonSeg;
  const double dtd = Td / N;
  Real S = (Real)S0d, v = (Real)v0d;
  PathSum avg;
  for (int seg = 0; seg < nseg; ++seg) {
    ck[(size_t)(2 * seg) * stride] = S;
    ck[(size_t)(2 * seg + 1) * stride] = v;
#pragma unroll
    for (int j = 0; j < kHestonSeg; j += 2) {
      const int k = seg * kHestonSeg + j;
      if (k < N) heston_philox_pair<Real, false>(c, S, v, avg, U4{p, (uint32_t)(k >> 1), c2, w}, k0, k1, k + 1 < N);
    }
  }
  g[0] = g[1] = g[2] = 0.0;
  const double x = avg.value() / (double)N - Kd;
  if (!(x > 0.0)) return 0.0;
  const double disc = exp(-r * Td);
  const Real abar = (Real)(disc / N);
  Real U = 0, vbar = 0, ebar_sum = 0;
  for (int seg = nseg - 1; seg >= 0; --seg) {
    // per step of the segment: the spot after it and the step's derivatives
    // de/dv, dv'/dv at its starting variance (formed in the recompute, so the
    // reverse step is three multiply-adds)
    Real Sa[kHestonSeg], de[kHestonSeg], dv[kHestonSeg];
    Real Sr = ck[(size_t)(2 * seg) * stride];
    Real vr = ck[(size_t)(2 * seg + 1) * stride];
#pragma unroll
    for (int j = 0; j < kHestonSeg; j += 2) {
      const int k = seg * kHestonSeg + j;
      if (k < N) {
        // recompute exactly as the forward sweep stepped (heston_philox_pair)
        Real ra, ca, sa, rb, cb, sb;  // two steps' (squared radius, cos, sin) or (1, z1, z2)
        if constexpr (sizeof(Real) == 4) {
          const U4 q = philox4x32_10(U4{p, (uint32_t)(k >> 1), c2, w}, k0, k1);
          box_muller_polar(q.x, q.y, ra, ca, sa);
          box_muller_polar(q.z, q.w, rb, cb, sb);
        } else {
          normals4<Real>(U4{p, (uint32_t)(k >> 1), c2, w}, k0, k1, ca, sa, cb, sb);
          ra = rb = Real(1);
        }
        heston_step_derivs(c, Sr, vr, ra, ca, sa, de[j], dv[j]);
        Sa[j] = Sr;
        if (k + 1 < N) {
          heston_step_derivs(c, Sr, vr, rb, cb, sb, de[j + 1], dv[j + 1]);
          Sa[j + 1] = Sr;
        }
      }
    }
#pragma unroll
    for (int j = kHestonSeg - 1; j >= 0; --j) {
      const int k = seg * kHestonSeg + j;
      if (k < N) {
        U += abar * Sa[j];
        ebar_sum += U;
        vbar = vbar * dv[j] + U * de[j];
      }
    }
  }
  g[0] = (double)U / S0d;
  g[1] = (double)vbar * 2.0 * sqrt(v0d);
  g[2] = -Td * disc * x + dtd * (double)ebar_sum;
  return x;
}

// Forward (tangent) Greeks of one Heston-Asian path: the three derivatives
// the adjoint above computes, carried forward with the path instead of
// recovered by a reverse sweep.  With vp = max(v, 0) and, on the vp > 0
// branch, de/dv = sqrt(dt) z1 / (2 sqrt(vp)) - dt/2 and
// dv'/dv = 1 - kappa dt + xi sqrt(dt) zc / (2 sqrt(vp)) (else 0 and 1):
//   dS_{k+1} = dS_k e^{e_k} + S_{k+1} de/dv dv_k,  dv_{k+1} = dv_k dv'/dv,
// seeded dS_0 = 0, dv_0 = 2 sqrt(v0) (the derivative with respect to
// sqrt(v0)); the spot is linear in S0 (dS_k/dS0 = S_k / S0) and its log is
// linear in r (dS_k/dr = S_k k dt), so those two need no recursion.  Then
//   dP/dS0 = e^{-rT} A / S0,  dP/dsqrt(v0) = e^{-rT} (1/N) sum_k dS_k,
//   dP/dr = e^{-rT} (1/N) sum_k S_k k dt - T P   (A > K; all zero else).
// The time-weighted sum comes from the running sums A_j = sum_{i<=j} S_i the
// average already forms: sum_{k=1..N} k S_k = (N + 1) A_N - sum_{j=1..N} A_j,
// so a step adds its running sum (in FP64) instead of carrying k dt and a
// product.
// The path itself is stepped exactly as heston_philox_pair steps it, so the
// price is bit for bit the forward kernel's.  Three Greeks cost a reciprocal
// and ~12 instructions per step here, against a segment recompute and a
// reverse step for the adjoint.  Returns the undiscounted payoff; g as the
// adjoint's.
template <typename Real>
struct HestonTangent {
  Real dS = 0, dv = 0;  // dS_k/dsqrt(v0), dv_k/dsqrt(v0)
  PathSum dA, B;        // sum_k dS_k, sum_j A_j (the running sums of the spots)
};

// One step of heston_step_polar (FP32, polar normals) or heston_step (FP64,
// normals z1 = cs, z2 = sn, r2 = 1) with the tangent; the path arithmetic is
// the same instructions as those steps'.  Returns the new spot's tangent
// terms through tg; the caller adds the pair sums.
template <typename Real>
__device__ __forceinline__ void heston_step_tangent(const HestonConsts<Real>& c, Real& S, Real& v, Real r2, Real cs,
                                                    Real sn, HestonTangent<Real>& tg, Real& dS_out) {
#pragma clang fp contract(off)
  const Real vp = positive_part(v);
  Real q, u;  // q: the step's sqrt(vp) times the radius; u = radius / (2 sqrt(vp))
  if constexpr (sizeof(Real) == 4) {
    q = fsqrt(vp * r2);
    u = vp > Real(0) ? Real(0.5) * q * frcp(vp) : Real(0);
  } else {
    q = fsqrt(vp);  // r2 = 1: the normals carry their radius
    u = vp > Real(0) ? Real(0.5) / q : Real(0);
  }
  const Real e = ffma(c.c_sdt * q, cs, ffma(-c.c_hdt, vp, c.c_rdt));
  const Real zz = ffma(c.xr, cs, c.xs * sn);
  const Real growth = fexp2(e);
  const Real Sn = S * growth;
  const Real dedv = vp > Real(0) ? ffma(c.sdt_n * u, cs, -c.hdt_n) : Real(0);
  const Real dvdv = vp > Real(0) ? ffma(zz, u, c.one_kdt) : Real(1);
  tg.dS = ffma(tg.dS, growth, Sn * (dedv * tg.dv));
  tg.dv = tg.dv * dvdv;
  dS_out = tg.dS;
  S = Sn;
  v = ffma(q, zz, ffma(c.kdt, c.theta - vp, v));
}

template <typename Real>
__device__ __forceinline__ double heston_tangent_path(const HestonConsts<Real>& c, double S0d, double Kd, double Td,
                                                      double v0d, double r, int N, uint32_t p, uint32_t c2,
                                                      uint32_t w, uint32_t k0, uint32_t k1, double (&g)[3]) {
#pragma clang fp contract(off)
  Real S = (Real)S0d, v = (Real)v0d;
  PathSum avg;
  HestonTangent<Real> tg;
  tg.dv = (Real)(2.0 * sqrt(v0d));
  for (int k = 0; k < N; k += 2) {
    const bool second = k + 1 < N;
    Real ra, ca, sa, rb, cb, sb;  // two steps' (squared radius, cos, sin) or (1, z1, z2)
    if constexpr (sizeof(Real) == 4) {
      const U4 q = philox4x32_10(U4{p, (uint32_t)(k >> 1), c2, w}, k0, k1);
      box_muller_polar(q.x, q.y, ra, ca, sa);
      box_muller_polar(q.z, q.w, rb, cb, sb);
    } else {
      normals4<Real>(U4{p, (uint32_t)(k >> 1), c2, w}, k0, k1, ca, sa, cb, sb);
      ra = rb = Real(1);
    }
    Real d0, d1;
    heston_step_tangent(c, S, v, ra, ca, sa, tg, d0);
    const double Aprev = avg.value();
    if (second) {
      const Real sa0 = S;
      heston_step_tangent(c, S, v, rb, cb, sb, tg, d1);
      avg.add(sa0 + S);
      tg.dA.add(d0 + d1);
      // A_{k+1} + A_{k+2} = 2 A_k + 2 S_{k+1} + S_{k+2}
      tg.B.add(fma(2.0, Aprev, (double)ffma(Real(2), sa0, S)));
    } else {
      avg.add(S);
      tg.dA.add(d0);
      tg.B.add(Aprev + (double)S);
    }
  }
  g[0] = g[1] = g[2] = 0.0;
  const double A = avg.value() / (double)N;
  const double x = A - Kd;
  if (!(x > 0.0)) return 0.0;
  const double disc = exp(-r * Td);
  g[0] = disc * A / S0d;
  g[1] = disc * tg.dA.value() / (double)N;
  const double St = Td / N * ((double)(N + 1) * avg.value() - tg.B.value());  // sum_k S_k k dt
  g[2] = disc * St / (double)N - Td * disc * x;
  return x;
}

// ---------------------------------------------------------------------------
// GBM European call (MonteCarloInsideBlockingDP.cpp:239-257, log2 domain)
// ---------------------------------------------------------------------------
template <typename Real>
struct GbmConsts {
  Real mu, vol, Y, Z, sqrtT, sigT;
  __device__ GbmConsts(double S, double X, double T, double r, double sigma) {
    const double sqrtTd = sqrt(T);
    mu = (Real)(kPathLog2e * (r - 0.5 * sigma * sigma) * T);
    vol = (Real)(kPathLog2e * sigma * sqrtTd);
    Y = (Real)S;
    Z = (Real)X;
    sqrtT = (Real)sqrtTd;
    sigT = (Real)(sigma * T);
  }
};

// Terminal value of one path and, when in the money, the pathwise
// derivative kernels: returns c = S_T - K; st = S_T; dlog = dlnS_T/dsigma.
template <typename Real>
__device__ __forceinline__ Real gbm_path(const GbmConsts<Real>& g, Real z, Real& st, Real& dlog) {
  st = g.Y * fexp2(g.mu + g.vol * z);
  dlog = g.sqrtT * z - g.sigT;
  return st - g.Z;
}

}  // namespace fsmc

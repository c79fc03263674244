// LIBOR market model swaption portfolio with pathwise adjoint deltas
// (the liborSwaptionGreeks sample, after Giles).  One path per lane:
//
//  forward   path_calc_b1, liborSwaptionGreeks.cpp:34-82 -- 40 log-Euler steps
//            of the 80 forward rates under the spot measure, rates in
//            registers; every updated rate is written to the trajectory
//            buffer (triangular rows, [slot][path] so that a wavefront's
//            stores and loads are coalesced);
//  portfolio portfolio_b, :120-211 -- 15 payer swaptions and the reverse
//            sweep of the annuity/discount recursion (in registers);
//  adjoint   path_calc_b2, :86-117 -- 40 reverse steps reading the
//            trajectory back one row at a time (the whole row is prefetched
//            before the step so its latency is paid once per row).
//
// Each path writes its value and 80 deltas to a [81][path] buffer; a second
// kernel reduces them in FP64 (mean and standard error per quantity).
//
// FSMC_QUIRK_LIBOR_INPLACE_ADJ reproduces the sample computing L_b in place
// over L (:173-184): B_b[m-1] is then divided by 1 + delta*L_b[n] instead of
// 1 + delta*L[n].
#include "fsmc_device.hpp"
#include "fsmc_host.hpp"

namespace fsmc {
namespace {

constexpr int NN = 80;    // forward rates           (params.h:27)
constexpr int NMAT = 40;  // simulation steps        (params.h:28)
constexpr int NOPT = 15;  // swaptions               (params.h:30)
constexpr int kBlock = 128;

// Portfolio of params.h:38-40.  The sample initialises its REAL swap rates
// from float literals, so the FP64 build sees the float-rounded values.
__host__ __device__ constexpr int mat_of(int j) {
  return j < 3 ? 4 : j < 6 ? 8 : j < 9 ? 20 : j < 12 ? 28 : 40;
}
__host__ __device__ constexpr float rate_of(int j) { return (j % 3) == 0 ? .045f : (j % 3) == 1 ? .05f : .055f; }

// Trajectory row r (1..40) holds L(r, i) for i = r..79; its first slot:
__host__ __device__ constexpr int row_off(int r) { return 80 * (r - 1) - (r - 1) * r / 2; }
constexpr int kSlots = row_off(NMAT + 1);  // 2380
static_assert(kSlots == 2380, "trajectory size");

struct LiborArgs {
  int n_paths;
  double L0, lambda;
  const double* z;  // buffer mode: [n_paths][40] (liborSwaptionGreeks.cpp:386, z[path*NMAT + n])
  uint32_t k0, k1, w;
  void* traj;  // Real [kSlots][n_paths]
  void* out;   // Real [81][n_paths]: value, then dV/dL0_i
};

template <typename Real, bool kBuffer>
__device__ __forceinline__ Real normal_at(const LiborArgs& a, int p, int n, Real (&zq)[4]) {
  if (kBuffer) return (Real)a.z[(size_t)p * NMAT + n];
  if ((n & 3) == 0) normals4<Real>(U4{(uint32_t)p, (uint32_t)(n >> 2), 0u, a.w}, a.k0, a.k1, zq[0], zq[1], zq[2], zq[3]);
  const int j = n & 3;
  return j == 0 ? zq[0] : j == 1 ? zq[1] : j == 2 ? zq[2] : zq[3];
}

// portfolio_b (liborSwaptionGreeks.cpp:120-211) on the final rates: returns
// the discounted portfolio value and hands each dV/dL(40, n) to put(n, lb).
// The sample keeps the bond prices B[m] and the adjoints B_b[m], S_b[m] in
// arrays of 40; here the adjoints are running scalars (S_b and B_b of date m
// only feed date m - 1), the swaption payoffs' contributions to them are five
// pairs of scalars (one per maturity), and the reverse sweep recomputes B[m]
// forward from a checkpoint every 8 dates -- the same divisions in the same
// order as the forward sweep, so every value equals the array formulation's
// bit for bit while the section needs ~20 registers besides the rates and
// adjoints.  The discount b to date 40 is taken first (the same divisions in
// the same order as the sample's, :200-206), so each adjoint leaves in its
// final, discounted form the moment it is made: no array of unscaled
// adjoints waits for b.  The phases are kept apart by scheduling barriers, so
// the compiler cannot hoist the next phase's loads into this one's registers.
// lo(n) returns rate n < 40 (registers or LDS, as the caller keeps them).
constexpr int kSeg = 8;
constexpr int kMats = 5;  // distinct maturities of params.h:38
__host__ __device__ constexpr int mat_slot(int m) {  // m = maturity - 1 -> its slot, or -1
  return m == 3 ? 0 : m == 7 ? 1 : m == 19 ? 2 : m == 27 ? 3 : m == 39 ? 4 : -1;
}
template <typename Real, bool kInplace, typename Lo, typename Put>
__device__ __forceinline__ Real portfolio_adjoint(const Real (&L)[NN], Lo lo, Put put) {
  using O = MathOps<Real>;
  const Real delta = Real(0.25);
  Real disc = 1;  // 1 / prod_{n < 40} (1 + delta L_n)
#pragma unroll
  for (int n = 0; n < NMAT; ++n) disc = O::div(disc, Real(1) + delta * lo(n));
  // disc is made here (not sunk to its first use with its 40 operands held
  // live), and lo(n) is read again at the end rather than kept
  asm volatile("" : "+v"(disc)::"memory");
  __builtin_amdgcn_sched_barrier(0);
  Real bck[NMAT / kSeg];  // B at the start of each segment (B[-1] = 1 for the first)
  Real injS[kMats], injB[kMats];
  Real b = 1, s = 0, val = 0;
#pragma unroll
  for (int m = 0; m < NMAT; ++m) {
    if (m % kSeg == 0) bck[m / kSeg] = b;
    b = O::div(b, Real(1) + delta * L[m + NMAT]);
    s = s + delta * b;
    if (mat_slot(m) >= 0) {
      Real sb = 0, bb = 0;
#pragma unroll
      for (int j = 0; j < NOPT; ++j) {
        if (mat_of(j) - 1 == m) {
          const Real K = (Real)rate_of(j);
          const Real sw = b + K * s - Real(1);
          if (sw < Real(0)) {
            val += Real(-100) * sw;
            sb += Real(-100) * K;
            bb += Real(-100);
          }
        }
      }
      injS[mat_slot(m)] = sb;
      injB[mat_slot(m)] = bb;
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  Real Sb = 0, Bb = 0;  // S_b[m], B_b[m] of the date being processed
#pragma unroll
  for (int g = NMAT / kSeg - 1; g >= 0; --g) {
    Real B[kSeg];  // B[g kSeg .. g kSeg + kSeg - 1], recomputed as in the forward sweep
    Real Lq[kSeg];  // opaque copies of the segment's rates: the forward sweep's 40
                    // reciprocals are recomputed here, not kept live for reuse
#pragma unroll
    for (int q = 0; q < kSeg; ++q) {
      Lq[q] = L[g * kSeg + q + NMAT];
      asm volatile("" : "+v"(Lq[q]));
    }
    Real bb0 = bck[g];
#pragma unroll
    for (int q = 0; q < kSeg; ++q) {
      bb0 = O::div(bb0, Real(1) + delta * Lq[q]);
      B[q] = bb0;
    }
#pragma unroll
    for (int q = kSeg - 1; q >= 0; --q) {
      const int m = g * kSeg + q, n = m + NMAT;
      if (mat_slot(m) >= 0) {
        Sb += injS[mat_slot(m)];
        Bb += injB[mat_slot(m)];
      }
      Bb += delta * Sb;
      const Real dn = Real(1) + delta * Lq[q];
      const Real lb = O::div(-Bb * B[q] * delta, dn);
      put(n, disc * lb);
      if (m > 0) Bb = O::div(Bb, kInplace ? Real(1) + delta * lb : dn);
    }
  }
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
  val = disc * val;
#pragma unroll
  for (int n = 0; n < NMAT; ++n) put(n, O::div(-val * delta, Real(1) + delta * lo(n)));
  return val;
}

template <typename Real, bool kBuffer, bool kInplace>
__global__ __launch_bounds__(kBlock) void libor_path_kernel(LiborArgs a) {
  using O = MathOps<Real>;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= a.n_paths) return;  // no workgroup synchronisation below
  const size_t P = (size_t)a.n_paths;
  Real* traj = (Real*)a.traj + p;
  const Real delta = Real(0.25), sqrt_delta = Real(0.5);
  const Real lam = (Real)a.lambda;
  const Real con1 = delta * lam;
  const Real L0 = (Real)a.L0;

  // ---- forward: path_calc_b1 ------------------------------------------------
  Real L[NN];
#pragma unroll
  for (int i = 0; i < NN; ++i) L[i] = L0;
  Real zq[4] = {0, 0, 0, 0};
  for (int n = 0; n < NMAT; ++n) {
    const Real sqez = sqrt_delta * normal_at<Real, kBuffer>(a, p, n, zq);
    // vrat = exp(con1*v + lam*(sqez - con1/2)), exponent pre-scaled for exp2
    const Real ce = con1 * O::kE;
    const Real c2 = lam * (sqez - Real(0.5) * con1) * O::kE;
    Real v = 0;
    const int rb = row_off(n + 1) - (n + 1);  // slot of entry i is rb + i (i > n)
#pragma unroll
    for (int i = 1; i < NN; ++i) {
      if (i > n) {
        const Real Li = L[i];
        v += con1 * O::div(Li, Real(1) + delta * Li);
        L[i] = Li * O::ex(ce * v + c2);
        traj[(size_t)(rb + i) * P] = L[i];
      }
    }
  }

  // ---- portfolio_b ---------------------------------------------------------------
  Real Lb[NN];
  const Real val = portfolio_adjoint<Real, kInplace>(
      L, [&](int n) { return L[n]; }, [&](int n, Real lb) { Lb[n] = lb; });

  // ---- adjoint: path_calc_b2 (L holds row n+1, Lb the adjoints) ---------------------
  for (int n = NMAT - 1; n >= 0; --n) {
    Real R[NN];  // row n, prefetched: entries i > n (clamped addresses below)
    if (n > 0) {
      // row n holds entries i >= n at slots row_off(n) + i - n; entries
      // i < n are not needed and read the (valid) slot of i = n instead
      const int rb = row_off(n) - n;
#pragma unroll
      for (int i = 1; i < NN; ++i) R[i] = traj[(size_t)(rb + (i > n ? i : n)) * P];
    } else {
#pragma unroll
      for (int i = 1; i < NN; ++i) R[i] = L0;
    }
    Real v1 = 0;
#pragma unroll
    for (int i = NN - 1; i >= 1; --i) {
      if (i > n) {
        const Real Ln = L[i], Lc = R[i];
        v1 += lam * Ln * Lb[i];
        const Real faci = O::div(delta, Real(1) + delta * Lc);
        Lb[i] = Lb[i] * O::div(Ln, Lc) + v1 * lam * faci * faci;
        L[i] = Lc;
      }
    }
  }

  Real* out = (Real*)a.out + p;
  out[0] = val;
#pragma unroll
  for (int i = 0; i < NN; ++i) out[(size_t)(i + 1) * P] = Lb[i];
}

// Largest lambda for which the FP32 kernel rebuilds rows instead of storing
// them: the third-order series below is then within 1.5e-8 of the exact
// inverse for every rate (2.3e-5 at lambda = 0.75).
constexpr double kRebuildMaxLambda = 0.3;

// Checkpoints of the FP32 rebuild kernel (kCkpt, fsmc_libor_params::
// checkpoint_every = 8): the forward sweep stores rows 8, 16, 24 and 32
// (entries i >= row), and the reverse sweep replaces the rebuilt row by the
// stored one at the top of each segment, so a segment's rows are recomputed
// from the exact row above it and the series' truncation never accumulates
// over more than kCkptEvery steps.  [slot][path] layout; row c starts at slot
// ckpt_base(c).  Measured on the 262,144-path batch: 0.64 ms without
// checkpoints, 0.71 ms with them (the 251 MB of stores cost ~20 us, the
// reloads, which every wave stalls on at nearly the same step, ~45 us; rows
// staggered over the waves measured 0.74 ms, and reading the checkpoint in
// place of the step's inversion spilled the kernel).
constexpr int kCkptEvery = 8;
__host__ __device__ constexpr bool ckpt_row(int c) { return c < NMAT && c % kCkptEvery == 0 && c > 0; }
__host__ __device__ constexpr int ckpt_base(int c) {  // rows 8j below c take 80 - 8j slots each
  return (c / kCkptEvery - 1) * NN - 4 * (c / kCkptEvery - 1) * (c / kCkptEvery) + 0 * c;
}
constexpr int kCkptSlots = 240;  // rows 8, 16, 24, 32: 72 + 64 + 56 + 48
static_assert(ckpt_base(8) == 0 && ckpt_base(16) == 72 && ckpt_base(24) == 136 && ckpt_base(32) == 192,
              "checkpoint slots");
static_assert(ckpt_base(32) + NN - 32 == kCkptSlots, "checkpoint slots");

// FP32 variant without the trajectory buffer.  The reverse sweep rebuilds row
// n from row n+1 instead of reading it back: for i = n+1, n+2, ... the forward
// map L(n+1,i) = x exp(con1 (v(i-1) + con1 x / (1 + delta x)) + c2(n)) with
// x = L(n,i) is inverted through the fixed point
//   x = w exp(-a t(x)),  t(x) = x / (1 + delta x),  a = con1^2,
//   w = L(n+1,i) exp(-(con1 v(i-1) + c2(n))).
// With s = a t(w) and r0 = 1 / (1 + delta w), its solution is the series
//   x = w (1 - s + (r0 + 1/2) s^2 - (2 r0^2 + r0/2 + 1/6) s^3 + O(s^4))
// (at delta = 0 this is Lambert's W(aw)/a = w (1 - s + 3/2 s^2 - 8/3 s^3)).
// s < a / delta = delta lambda^2 for any rate, so the truncation stays below
// FP32 rounding for lambda <= kRebuildMaxLambda; larger lambdas take the
// stored-trajectory kernel.  The adjoint then runs in the same ascending
// order: its running sum v1(i) = sum_{k >= i} lambda L(n+1,k) Lb(k) is taken
// as the row total minus the prefix below i.  Rows 8, 16, 24 and 32 are
// checkpointed in the forward sweep (kCkptEvery) and restore the rebuild at
// the top of each segment; the other HBM traffic is the [81][path] outputs.
template <bool kBuffer, bool kInplace, bool kCkpt>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void libor_path_rc_kernel(
    LiborArgs a) {  // three waves per SIMD: <= 168 registers per lane; the sweeps do not spill (the portfolio
                    // section around them does, once per path)
  using O = MathOps<float>;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= a.n_paths) return;  // no workgroup synchronisation below
  const size_t P = (size_t)a.n_paths;
  const float delta = 0.25f, sqrt_delta = 0.5f;
  const float lam = (float)a.lambda;
  const float con1 = delta * lam;
  const float kk = con1 * con1 * O::kE;  // log2 exponent per unit of sum L / (1 + delta L)
  const float L0 = (float)a.L0;

  // ---- forward: path_calc_b1 (no stores) ---------------------------------------
  // vs = v / con1 = sum_k L_k / (1 + delta L_k), so the log2 exponent of the
  // rate ratio is kk vs + c2 with kk = con1^2 log2(e).
  float L[NN];
#pragma unroll
  for (int i = 0; i < NN; ++i) L[i] = L0;
  float zq[4] = {0.f, 0.f, 0.f, 0.f};
  float* ckpt = (float*)a.traj + p;
  for (int n = 0; n < NMAT; ++n) {
    const float sqez = sqrt_delta * normal_at<float, kBuffer>(a, p, n, zq);
    const float c2 = lam * (sqez - 0.5f * con1) * O::kE;
    float vs = 0.f;
#pragma unroll
    for (int i = 1; i < NN; ++i) {
      if (i > n) {
        const float Li = L[i];
        vs = __builtin_fmaf(Li, __builtin_amdgcn_rcpf(__builtin_fmaf(delta, Li, 1.f)), vs);
        L[i] = Li * O::ex(__builtin_fmaf(kk, vs, c2));
      }
    }
    if (kCkpt && ckpt_row(n + 1)) {  // uniform: row n + 1 is a checkpoint, entries i >= n + 1
      const int cb = ckpt_base(n + 1) - (n + 1);  // slot of entry i is cb + i
#pragma unroll
      for (int i = 1; i < NN; ++i)
        if (i > n) ckpt[(size_t)(cb + i) * P] = L[i];
    }
  }

  // The reverse sweep carries M(i) = Lb(i) L(n+1,i) instead of Lb(i): the
  // reference's update Lb(n,i) = Lb(n+1,i) L(n+1,i) / L(n,i) + v1 lambda faci^2
  // becomes M(n,i) = M(n+1,i) + L(n,i) lambda v1 faci^2, with no division, and
  // the row total S = sum_{k>n} L(n+1,k) Lb(k) is the plain sum of M.  Every
  // rate starts at L0, so Lb(0,i) = M(0,i) / L0.
  // Rates 1..39 move to LDS for the portfolio and the reverse sweep (each
  // thread its own column, no barrier needed): with only rates 0 and
  // 40..79 and the adjoints in registers the kernel fits three waves per
  // SIMD without spilling.
  __shared__ float Llo[NMAT][kBlock];
#pragma unroll
  for (int i = 1; i < NMAT; ++i) Llo[i][threadIdx.x] = L[i];
  // compiler barrier (no instruction): the rates are read back from LDS
  // below instead of being kept live in registers
  asm volatile("" ::: "memory");
  float M[NN];
  const float L0r = L[0];
  auto lo = [&](int n) { return n == 0 ? L0r : Llo[n][threadIdx.x]; };
  const float val = portfolio_adjoint<float, kInplace>(
      L, lo, [&](int n, float lb) { M[n] = n == 0 ? lb : lb * (n < NMAT ? lo(n) : L[n]); });

  // ---- adjoint: path_calc_b2, rows rebuilt in registers ------------------------
  // v1(i) lambda faci^2 = con1^2 Q / (1 + delta L(n,i))^2, with Q the row total
  // of M from i up.  The sweep keeps its rates scaled by sig = con1^2 (a floor
  // keeps sig a normal number as lambda -> 0, where every con1^2 term
  // vanishes): with x' = sig x the series variable is s = x' r0 and the
  // adjoint's factor con1^2 x r^2 is x' r^2, one multiply each fewer, and the
  // running sum vs' = sig vs enters the exponent as kk / sig.
  const float con1sq = con1 * con1;
  const float sig = con1sq > 1e-30f ? con1sq : 1e-30f;
  const float kks = kk / sig, dls = delta / sig;
#pragma unroll
  for (int i = 1; i < NN; ++i) {
    if (i < NMAT)
      Llo[i][threadIdx.x] *= sig;
    else
      L[i] *= sig;
  }
  for (int n = NMAT - 1; n >= 0; --n) {
    if (kCkpt && ckpt_row(n + 1)) {  // uniform: the stored row n + 1 replaces the rebuilt one
      const int cb = ckpt_base(n + 1) - (n + 1);
#pragma unroll
      for (int i = 1; i < NN; ++i) {
        if (i > n) {
          const float y = sig * ckpt[(size_t)(cb + i) * P];
          if (i < NMAT)
            Llo[i][threadIdx.x] = y;
          else
            L[i] = y;
        }
      }
    }
    float z;
    if (kBuffer) {
      z = (float)a.z[(size_t)p * NMAT + n];
    } else {
      // regenerated per step (one Philox call per step, not per four): keeps
      // the four normals out of the registers live across the row loops
      float zr[4];
      normals4<float>(U4{(uint32_t)p, (uint32_t)(n >> 2), 0u, a.w}, a.k0, a.k1, zr[0], zr[1], zr[2], zr[3]);
      const int j = n & 3;
      z = j == 0 ? zr[0] : j == 1 ? zr[1] : j == 2 ? zr[2] : zr[3];
    }
    const float c2 = lam * (sqrt_delta * z - 0.5f * con1) * O::kE;
    float Q = 0.f;
#pragma unroll
    for (int i = 1; i < NN; ++i)
      if (i > n) Q += M[i];
    float vs = 0.f;
#pragma unroll
    for (int i = 1; i < NN; ++i) {
      if (i > n) {
        const float y = i < NMAT ? Llo[i][threadIdx.x] : L[i];
        const float w = y * O::ex(-__builtin_fmaf(kks, vs, c2));
        const float r0 = __builtin_amdgcn_rcpf(__builtin_fmaf(dls, w, 1.f));
        // x = w (1 + e), e = -s + c2 s^2 + c3 s^3, s = con1^2 w / (1 + delta w)
        const float s = w * r0;
        const float c3 = __builtin_fmaf(r0, __builtin_fmaf(-2.f, r0, -0.5f), -1.f / 6.f);
        const float h = __builtin_fmaf(s, __builtin_fmaf(s, c3, r0 + 0.5f), -1.f);
        const float x = __builtin_fmaf(w, s * h, w);
        const float r = __builtin_amdgcn_rcpf(__builtin_fmaf(dls, x, 1.f));
        const float xr = x * r;
        vs += xr;
        const float m = M[i];
        M[i] = __builtin_fmaf(xr * r, Q, m);
        Q -= m;
        if (i < NMAT)
          Llo[i][threadIdx.x] = x;
        else
          L[i] = x;
      }
    }
  }

  const float inv_L0 = (float)(1.0 / a.L0);
  float* out = (float*)a.out + p;
  out[0] = val;
  out[P] = M[0];
#pragma unroll
  for (int i = 1; i < NN; ++i) out[(size_t)(i + 1) * P] = M[i] * inv_L0;
}

// Sum and sum of squares of quantity q over a strided set of paths, in FP64.
template <typename Real>
__global__ __launch_bounds__(256) void libor_reduce_kernel(int n_paths, const Real* out, double* partials) {
  __shared__ double lds[2 * 4];
  const int q = blockIdx.y;
  const Real* src = out + (size_t)q * n_paths;
  double s[2] = {0.0, 0.0};
  for (int p = blockIdx.x * 256 + threadIdx.x; p < n_paths; p += gridDim.x * 256) {
    const double x = (double)src[p];
    s[0] += x;
    s[1] += x * x;
  }
  block_sum<2>(s, lds);
  if (threadIdx.x == 0) {
    double* d = partials + ((size_t)q * gridDim.x + blockIdx.x) * 2;
    d[0] = s[0];
    d[1] = s[1];
  }
}

// One wavefront per output q (workgroup of 64; q = 0 the value, q = 1..NN
// the deltas): the chunk partials combined by wave_combine, mean and standard
// error written straight into the caller's value[2] / deltas[2 NN].
__global__ __launch_bounds__(64) void libor_finalize_kernel(int n_paths, int nchunks, const double* partials,
                                                           double* value, double* deltas) {
  const int q = blockIdx.x;
  double s[2];
  wave_combine<2>(partials + (size_t)q * nchunks * 2, nchunks, 2, s);
  if (threadIdx.x != 0) return;
  const double N = (double)n_paths;
  const double mean = s[0] / N;
  const double var = (N * s[1] - s[0] * s[0]) / (N * (N - 1.0));
  const double se = sqrt(var > 0.0 ? var : 0.0) / sqrt(N);
  double* dst = q == 0 ? value : deltas + (q - 1);
  dst[0] = mean;
  dst[q == 0 ? 1 : NN] = se;
}

// Normals of the LIBOR Philox layout written out as a [n_paths][40] buffer,
// scaled to N(mean, stddev): the device side of the generateRand seam
// (mklRng.h:9).  A buffer-mode run on this z equals the Philox-mode run.
__global__ void libor_normals_kernel(int n_paths, double mean, double stddev, uint32_t k0, uint32_t k1, uint32_t w,
                                     double* z) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // one Philox call each
  const int p = t / (NMAT / 4), q = t % (NMAT / 4);
  if (p >= n_paths) return;
  double n[4];
  normals4<double>(U4{(uint32_t)p, (uint32_t)q, 0u, w}, k0, k1, n[0], n[1], n[2], n[3]);
  double* dst = z + (size_t)p * NMAT + 4 * q;
#pragma unroll
  for (int j = 0; j < 4; ++j) dst[j] = mean + stddev * n[j];
}

constexpr int kReduceChunks = 64;

struct LiborPlan {
  size_t real, traj_bytes, out_bytes, part_bytes, ws;
};

bool stores_trajectory(const fsmc_libor_params* p, const fsmc_mode* m) {
  return m->precision == FSMC_FP64 || p->lambda > kRebuildMaxLambda;
}

LiborPlan libor_plan(const fsmc_libor_params* p, const fsmc_mode* m) {
  LiborPlan l;
  l.real = m->precision == FSMC_FP64 ? sizeof(double) : sizeof(float);
  const size_t P = (size_t)p->n_paths;
  // the FP32 kernel rebuilds the trajectory in registers up to kRebuildMaxLambda, from
  // checkpoint rows every kCkptEvery steps
  l.traj_bytes = stores_trajectory(p, m) ? align_up((size_t)kSlots * P * l.real, 256)
                 : p->checkpoint_every ? align_up((size_t)kCkptSlots * P * l.real, 256)
                                       : 0;
  l.out_bytes = align_up((size_t)(NN + 1) * P * l.real, 256);
  l.part_bytes = align_up((size_t)(NN + 1) * kReduceChunks * 2 * sizeof(double), 256);
  l.ws = l.traj_bytes + l.out_bytes + l.part_bytes;
  return l;
}

int libor_validate(const fsmc_libor_params* p, const fsmc_mode* m, bool mrg_allowed = false) {
  FSMC_REQUIRE(p != nullptr, "params is NULL");
  FSMC_TRY(validate_mode(m, true, mrg_allowed));
  FSMC_REQUIRE(p->n_paths >= 2 && p->n_paths <= (1 << 30), "n_paths %d outside [2, 2^30]", p->n_paths);
  FSMC_REQUIRE(p->L0 > 0.0, "L0 must be > 0");
  FSMC_REQUIRE(p->lambda >= 0.0, "lambda must be >= 0");
  FSMC_REQUIRE(p->checkpoint_every == 0 || p->checkpoint_every == kCkptEvery,
               "checkpoint_every %d: 0 or %d", p->checkpoint_every, kCkptEvery);
  return FSMC_OK;
}

template <typename Real, bool B, bool Q>
void launch(const LiborArgs& a, bool ckpt, hipStream_t s) {
  const dim3 grid((a.n_paths + kBlock - 1) / kBlock);
  if (sizeof(Real) == 4 && a.lambda <= kRebuildMaxLambda) {
    if (ckpt)
      hipLaunchKernelGGL((libor_path_rc_kernel<B, Q, true>), grid, dim3(kBlock), 0, s, a);
    else
      hipLaunchKernelGGL((libor_path_rc_kernel<B, Q, false>), grid, dim3(kBlock), 0, s, a);
  } else {
    hipLaunchKernelGGL((libor_path_kernel<Real, B, Q>), grid, dim3(kBlock), 0, s, a);
  }
}

template <typename Real>
void launch_all(const LiborArgs& a, bool buffer, bool q, bool ckpt, hipStream_t s) {
  if (buffer) q ? launch<Real, true, true>(a, ckpt, s) : launch<Real, true, false>(a, ckpt, s);
  else q ? launch<Real, false, true>(a, ckpt, s) : launch<Real, false, false>(a, ckpt, s);
}

}  // namespace
}  // namespace fsmc

using namespace fsmc;

extern "C" {

size_t fsmc_libor_workspace(const fsmc_libor_params* p, const fsmc_mode* m) {
  if (libor_validate(p, m) != FSMC_OK) return 0;
  return libor_plan(p, m).ws;
}

int fsmc_libor_swaption_greeks_async(const fsmc_libor_params* p, const fsmc_mode* m, const double* d_z,
                                     double* d_value, double* d_deltas, void* d_ws, size_t ws_bytes,
                                     void* stream) {
  FSMC_TRY(libor_validate(p, m));
  FSMC_REQUIRE(d_value && d_deltas, "NULL device pointer");
  const bool buffer = m->rng == FSMC_RNG_BUFFER;
  FSMC_REQUIRE(!buffer || d_z, "buffer mode needs z");
  const LiborPlan l = libor_plan(p, m);
  if (d_ws == nullptr || ws_bytes < l.ws) {
    set_error("workspace %zu bytes < required %zu", ws_bytes, l.ws);
    return FSMC_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* ws = (char*)d_ws;
  LiborArgs a;
  a.n_paths = p->n_paths;
  a.L0 = p->L0;
  a.lambda = p->lambda;
  a.z = d_z;
  a.k0 = (uint32_t)m->seed;
  a.k1 = (uint32_t)(m->seed >> 32);
  a.w = counter_w(kTagLibor, m->stream);
  a.traj = ws;
  a.out = ws + l.traj_bytes;
  double* part = (double*)(ws + l.traj_bytes + l.out_bytes);
  const bool q = (m->quirks & FSMC_QUIRK_LIBOR_INPLACE_ADJ) != 0;
  const bool f64 = m->precision == FSMC_FP64;
  const bool ckpt = p->checkpoint_every != 0;
  if (f64) launch_all<double>(a, buffer, q, ckpt, s);
  else launch_all<float>(a, buffer, q, ckpt, s);
  FSMC_HIP(hipGetLastError());
  if (f64)
    hipLaunchKernelGGL(libor_reduce_kernel<double>, dim3(kReduceChunks, NN + 1), dim3(256), 0, s, p->n_paths,
                       (const double*)a.out, part);
  else
    hipLaunchKernelGGL(libor_reduce_kernel<float>, dim3(kReduceChunks, NN + 1), dim3(256), 0, s, p->n_paths,
                       (const float*)a.out, part);
  FSMC_HIP(hipGetLastError());
  // value and its standard error, the 80 deltas and their standard errors
  hipLaunchKernelGGL(libor_finalize_kernel, dim3(NN + 1), dim3(64), 0, s, p->n_paths, kReduceChunks, part, d_value,
                     d_deltas);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

int fsmc_libor_normals(fsmc_ctx* ctx, int n_paths, const fsmc_mode* m, double mean, double stddev, double* z) {
  FSMC_REQUIRE(ctx != nullptr && z != nullptr, "NULL argument");
  FSMC_TRY(validate_mode(m, false));
  // one thread per four normals, indexed in int
  FSMC_REQUIRE(n_paths >= 1 && n_paths <= 200000000, "n_paths %d outside [1, 2e8]", n_paths);
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t bytes = (size_t)n_paths * NMAT * sizeof(double);
  FSMC_TRY(ctx_prepare(ctx, 0, bytes, 0, 0));
  hipStream_t s = ctx->stream;
  const int threads = n_paths * (NMAT / 4);
  hipLaunchKernelGGL(libor_normals_kernel, dim3((threads + 255) / 256), dim3(256), 0, s, n_paths, mean, stddev,
                     (uint32_t)m->seed, (uint32_t)(m->seed >> 32), counter_w(kTagLibor, m->stream),
                     (double*)ctx->normals.ptr);
  FSMC_HIP(hipGetLastError());
  FSMC_HIP(hipMemcpyAsync(z, ctx->normals.ptr, bytes, hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

int fsmc_libor_swaption_greeks(fsmc_ctx* ctx, const fsmc_libor_params* p, const fsmc_mode* m, const double* z,
                               double* value, double* deltas) {
  FSMC_REQUIRE(ctx != nullptr, "ctx is NULL");
  FSMC_TRY(libor_validate(p, m, true));
  FSMC_REQUIRE(value && deltas, "NULL host pointer");
  const bool mrg = m->rng == FSMC_RNG_MRG32K3A;
  const bool buffer = m->rng == FSMC_RNG_BUFFER || mrg;
  FSMC_REQUIRE(m->rng != FSMC_RNG_BUFFER || z, "buffer mode needs z");
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t zn = buffer ? (size_t)p->n_paths * NMAT : 0;
  FSMC_TRY(ctx_prepare(ctx, 0, zn * sizeof(double), (2 + 2 * NN) * sizeof(double), libor_plan(p, m).ws));
  hipStream_t s = ctx->stream;
  fsmc_mode mb = *m;  // the kernels read the generated stream as a buffer
  if (mrg) {
    // generateRand's layout z[40 path + step] is the stream order
    FSMC_TRY(mrg_gaussian_enqueue((uint32_t)m->seed, (uint64_t)m->stream * zn, (long long)zn, 0.0, 1.0,
                                  (double*)ctx->normals.ptr, s));
    mb.rng = FSMC_RNG_BUFFER;
  } else if (buffer) {
    FSMC_HIP(hipMemcpyAsync(ctx->normals.ptr, z, zn * sizeof(double), hipMemcpyHostToDevice, s));
  }
  double* dv = (double*)ctx->outputs.ptr;
  double* dd = dv + 2;
  int st = fsmc_libor_swaption_greeks_async(p, &mb, (const double*)ctx->normals.ptr, dv, dd, ctx->workspace.ptr,
                                            ctx->workspace.bytes, s);
  if (st != FSMC_OK) {
    (void)hipStreamSynchronize(s);
    return st;
  }
  FSMC_HIP(hipMemcpyAsync(value, dv, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipMemcpyAsync(deltas, dd, 2 * NN * sizeof(double), hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

}  // extern "C"

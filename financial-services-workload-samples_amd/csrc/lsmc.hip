// Longstaff-Schwartz American puts.
//
// Batched "correct" LSMC (many options, uneven step and path counts).  Paths
// are built backwards in time by the Brownian bridge, so no path matrix is
// stored: with W the driving Brownian motion at dates t_k = k dt,
//   W_N = sqrt(T) Z_{N-1},  W_k = k/(k+1) W_{k+1} + sqrt(dt k/(k+1)) Z_{k-1},
//   S_k = S0 exp((r - sigma^2/2) t_k + sigma W_k),
// exactly the law of the forward GBM path, Z_i of path p being component p & 3
// of Philox counter (p/4, i, option): one Philox call per four paths per date.  One state word per
// path (W at the current date) replaces the [date][path] matrix: a pass
// reads and rewrites it, so the 4 B per path-step matrix write and a second
// row read per pass disappear.
//   * lsmc_init_kernel: W_N, the terminal cash flow e^{-rT} (K - S_T)^+
//     (kept discounted to time 0) and the European sum;
//   * lsmc_pass_kernel, once per backward date: applies the exercise decision
//     of date d (spot from the stored W_d) with the coefficients fitted at d,
//     steps the bridge to W_{d-1} (stored back) and accumulates the
//     regression of the discounted cash flow on {1, x, x^2, x^3} over the
//     in-the-money paths of date d-1.  x = 1 + (S/K - 1)/(sigma sqrt(t)) is
//     an affine map of S/K (same fitted values as the plain monomials of S/K)
//     that keeps the Gram matrix well conditioned.  Each lane accumulates
//     the seven moments and four right-hand sides of its own paths in the
//     working precision, the workgroup combines them in FP64 (a 4x4 Gram matrix is too
//     small for matrix cores: a v_mfma_f32_4x4x1f32 formulation needs the
//     path's basis values spread over four lanes and cost ~117 instructions
//     per 64 paths against ~14 for per-lane FMAs).  Each workgroup writes its
//     partial sums;
//   * lsmc_solve_kernel, after each pass: one workgroup per option still
//     regressing combines that option's partials in FP64 (fixed order) and
//     solves the 4x4 system by Cholesky in FP64.  A separate launch rather
//     than a last-workgroup ticket: on this GPU every device-scope release
//     fence writes back the XCD's L2, and 35k such fences per pass cost
//     milliseconds, while the kernel boundary costs microseconds.
//   * lsmc_final_kernel: price = max(K - S0, mean discounted cash flow), from
//     the per-workgroup cash-flow sums the option's last pass (the decision
//     at date 1) leaves behind, so the cash flows are not read again.
// Options are sorted by step count (descending) so the options still active
// at backward pass j are a prefix of the workgroup table: pass j launches
// exactly their workgroups.  That is the streamed sweep; the resident sweep
// (below) runs the same arithmetic on the same normals in one cooperative
// launch with the paths' state in registers, and is the default when the
// batch fits the resident grid.
//
// Reference algorithm ("refcompat", MonteCarloAmericanOptions/main.cpp:115-279
// and utilities.cpp:42-123) for one option: path matrix, lagged quadratic
// regression with ridge 0.01, cofactor 3x3 inverse, no discounting; the
// per-step solve is a one-workgroup kernel after each step kernel.
#include <algorithm>
#include <numeric>
#include <vector>

#include "fsmc_device.hpp"
#include "fsmc_host.hpp"

namespace fsmc {
namespace {

constexpr int kBlock = 256;
constexpr int kPPT = 16;  // paths per thread per workgroup
constexpr int kChunk = kBlock * kPPT;
constexpr int kNS = 12;  // partial sums: m0..m6, c0..c3, spare
constexpr double kLog2e = 1.4426950408889634074;

struct OptMeta {
  long long cf_off;   // first cash flow and bridge state of the option
  int first_block, nblk, steps, paths;
  double S0, K, sigma, r, T;
};

struct LsmcArgs {
  const OptMeta* meta;
  const int4* table;  // workgroup -> (option, chunk, first cash-flow element, paths in the chunk)
  void* state;  // bridge state: W at the current date, one word per path (cf layout)
  void* cf;
  double* partials;    // [workgroup][kNS]
  const int* order;    // option indices sorted by steps (descending)
  double* beta;        // [2][option][5]: coefficients, valid flag (double-buffered by pass)
  double* eur;         // [workgroup] European payoff sums
  double* cfsum;       // [workgroup][2] sum and sum of squares of the final cash flows
  int n_options;
  int pass;
  uint32_t k0, k1, w;
};

template <typename Real>
__device__ __forceinline__ Real pick4(const Real (&z)[4], int j) {
  return j == 0 ? z[0] : j == 1 ? z[1] : j == 2 ? z[2] : z[3];
}

// ---------------------------------------------------------------------------
// Batched correct LSMC
// ---------------------------------------------------------------------------
// Normal i of paths 4g .. 4g+3 (path p takes component p & 3 of Philox
// counter (p / 4, i, option)): one call serves a thread's four consecutive
// paths at one date.
template <typename Real>
__device__ __forceinline__ void group_normals(int g, int i, uint32_t o, const LsmcArgs& a, Real (&z)[4]) {
  normals4<Real>(U4{(uint32_t)g, (uint32_t)i, o, a.w}, a.k0, a.k1, z[0], z[1], z[2], z[3]);
}

// A thread owns kPPT / 4 groups of four consecutive paths; per-option cash
// flow and state segments start on a multiple of 4 elements and are padded
// to one, so a live group is one aligned vector load / store.
constexpr int kGroups = kPPT / 4;
template <typename Real>
struct Quad {
  typedef Real V __attribute__((ext_vector_type(4)));
};
template <typename Real>
__device__ __forceinline__ void load4(const Real* src, Real (&v)[4]) {
  const typename Quad<Real>::V q = *reinterpret_cast<const typename Quad<Real>::V*>(src);
  v[0] = q[0];
  v[1] = q[1];
  v[2] = q[2];
  v[3] = q[3];
}
template <typename Real>
__device__ __forceinline__ void store4(Real* dst, const Real (&v)[4]) {
  typename Quad<Real>::V q;
  q[0] = v[0];
  q[1] = v[1];
  q[2] = v[2];
  q[3] = v[3];
  *reinterpret_cast<typename Quad<Real>::V*>(dst) = q;
}
__device__ __forceinline__ int group_first_path(int chunk, int h) {
  return chunk * kChunk + 4 * (h * kBlock + (int)threadIdx.x);
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void lsmc_init_kernel(LsmcArgs a) {
  using O = MathOps<Real>;
  __shared__ double lds[kBlock / 64];
  const int4 e = a.table[blockIdx.x];
  const OptMeta m = a.meta[e.x];
  const Real sqT = (Real)sqrt(m.T);
  const Real muT = (Real)((m.r - 0.5 * m.sigma * m.sigma) * m.T * O::kE), sg = (Real)(m.sigma * O::kE);
  const Real discT = (Real)exp(-m.r * m.T);
  const Real K = (Real)m.K, S0 = (Real)m.S0;
  Real* w = (Real*)a.state + m.cf_off;
  Real* cf = (Real*)a.cf + m.cf_off;
  double eu[1] = {0.0};
#pragma unroll
  for (int h = 0; h < kGroups; ++h) {
    const int p0 = group_first_path(e.y, h);
    if (p0 >= m.paths) continue;
    Real z[4];
    group_normals<Real>(p0 >> 2, m.steps - 1, (uint32_t)e.x, a, z);
    Real W[4], c[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      W[g] = sqT * z[g];
      const Real ST = S0 * O::ex(muT + sg * W[g]);
      c[g] = p0 + g < m.paths ? discT * (K - ST > Real(0) ? K - ST : Real(0)) : Real(0);
      eu[0] += (double)c[g];
    }
    store4(w + p0, W);  // padding entries are written too (never read as paths)
    store4(cf + p0, c);
  }
  block_sum<1>(eu, lds);
  if (threadIdx.x == 0) a.eur[blockIdx.x] = eu[0];
}

// Moments of one path: s[q] += x^q (q = 0..6), s[7 + q] += x^q y (q = 0..3)
// for an in-the-money path; other paths add exact zeros (branch-free).
constexpr int kNM = 11;
template <typename Real>
__device__ __forceinline__ void accum_moments(Real (&s)[kNM], bool itm, Real x, Real y) {
  Real xp = itm ? Real(1) : Real(0);
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    s[q] += xp;
    if (q < 4) s[7 + q] += xp * y;
    xp *= x;
  }
}

// Sums the NV-wide partials of workgroups [first, first + n) with every
// thread of the calling workgroup: thread t takes workgroups t, t + 256, ...
// in order, then the fixed block tree (so the result is reproducible).  The
// total lands in thread 0.
template <int NV>
__device__ __forceinline__ void combine_partials(const double* partials, int first, int n, double (&t)[NV],
                                                 double* lds) {
#pragma unroll
  for (int q = 0; q < NV; ++q) t[q] = 0.0;
  for (int b = threadIdx.x; b < n; b += kBlock) {
#pragma unroll
    for (int q = 0; q < NV; ++q) t[q] += partials[(size_t)(first + b) * NV + q];
  }
  block_sum<NV>(t, lds);
}

// Solves the 4x4 normal equations from the option's summed moments s.
__device__ __forceinline__ double rsqrt_f64(double v) {
  double y = __builtin_amdgcn_rsq(v);
  const double hv = 0.5 * v;
#pragma unroll
  for (int it = 0; it < 2; ++it) y = fma(y, fma(-hv * y, y, 0.5), y);  // y += y (1/2 - v y^2 / 2)
  return y;
}
__device__ void lsmc_solve(const double (&s)[kNS], double* beta_out) {
  // Cholesky G = L L^T of G_ij = m_{i+j}; a pivot that is not clearly
  // positive (fewer than 4 ITM paths, or a singular basis) disables exercise
  // at this date.  The diagonal enters only through divisions, so its four
  // reciprocals are formed once and the remaining steps multiply: four
  // square roots and four divisions on the serial path instead of four and
  // fourteen (the solve sits on the critical path of every date).  Each
  // reciprocal square root is the hardware estimate refined by two Newton
  // steps (quadratic convergence from its ~2^-23 error to FP64 rounding),
  // not an IEEE square root followed by an IEEE division.  Fully unrolled so
  // L stays in registers.
  double L[4][4] = {{0}}, inv[4];
  bool ok = s[0] >= 4.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double v = s[i + j];
#pragma unroll
      for (int k = 0; k < j; ++k) v -= L[i][k] * L[j][k];
      if (i == j) {
        ok = ok && v > 1e-13 * s[2 * i];
        inv[i] = ok ? rsqrt_f64(v) : 1.0;
      } else {
        L[i][j] = v * inv[j];
      }
    }
  }
  double y[4], x[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double v = s[7 + i];
#pragma unroll
    for (int k = 0; k < i; ++k) v -= L[i][k] * y[k];
    y[i] = v * inv[i];
  }
#pragma unroll
  for (int i = 3; i >= 0; --i) {
    double v = y[i];
#pragma unroll
    for (int k = i + 1; k < 4; ++k) v -= L[k][i] * x[k];
    x[i] = v * inv[i];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) beta_out[i] = ok ? x[i] : 0.0;
  beta_out[4] = ok ? 1.0 : 0.0;
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void lsmc_pass_kernel(LsmcArgs a) {
  using O = MathOps<Real>;
  __shared__ double lds[kNS * (kBlock / 64)];
  // the table entry alone addresses the loads, so they leave while the
  // option's constants and coefficients are still in flight
  const int4 e = a.table[blockIdx.x];
  const int o = e.x;
  Real* wst = (Real*)a.state + e.z;
  Real* cf = (Real*)a.cf + e.z;
  Real cv[kGroups][4], wd[kGroups][4];
#pragma unroll
  for (int h = 0; h < kGroups; ++h) {
    const int l = 4 * (h * kBlock + (int)threadIdx.x);  // first path of the group within the chunk
    if (l < e.w) {
      load4(cf + l, cv[h]);
      load4(wst + l, wd[h]);
    } else {
#pragma unroll
      for (int g = 0; g < 4; ++g) cv[h][g] = wd[h][g] = Real(0);
    }
  }
  const OptMeta m = a.meta[o];
  const int d = m.steps - a.pass;  // decision date
  const int k = d - 1;             // regression date
  const double dt = m.T / m.steps;
  const double* bin = a.beta + (size_t)((a.pass & 1) * a.n_options + o) * 5;
  const bool decide = a.pass >= 1 && bin[4] != 0.0;
  const Real b0 = (Real)bin[0], b1 = (Real)bin[1], b2 = (Real)bin[2], b3 = (Real)bin[3];
  const Real K = (Real)m.K, invK = (Real)(1.0 / m.K), S0 = (Real)m.S0;
  const double mu = (m.r - 0.5 * m.sigma * m.sigma) * dt;
  const Real mu_d = (Real)(mu * d * O::kE), mu_k = (Real)(mu * k * O::kE), sg = (Real)(m.sigma * O::kE);
  const Real isw_d = (Real)(1.0 / (m.sigma * sqrt(d * dt))), disc_d = (Real)exp(-m.r * d * dt);
  const Real isw_k = k >= 1 ? (Real)(1.0 / (m.sigma * sqrt(k * dt))) : Real(0);
  // bridge from date d back to k = d - 1: W_k = k/d W_d + sqrt(dt k/d) Z_{k-1}
  const Real ba = (Real)((double)k / d), bb = (Real)sqrt(dt * k / d);
  // A thread's own paths (at most kPPT) are summed in the working precision.
  // In FP32 the thread sums (<= 16 terms) and the wave trees over them
  // (64 terms, 6 roundings) carry ~1e-6 relative rounding, independent
  // across waves, which are combined in FP64: the option's moments keep
  // ~1e-8.  FP64 runs are summed in FP64 throughout.
  Real sm[kNM];
#pragma unroll
  for (int q = 0; q < kNM; ++q) sm[q] = Real(0);
  double fin[2] = {0.0, 0.0};
#pragma unroll
  for (int h = 0; h < kGroups; ++h) {
    const int l0 = 4 * (h * kBlock + (int)threadIdx.x);
    if (l0 >= e.w) continue;
    const int p0 = e.y * kChunk + l0;  // path index of the group's first path
    Real z[4] = {0, 0, 0, 0}, wk[4];
    if (k >= 1) group_normals<Real>(p0 >> 2, k - 1, (uint32_t)o, a, z);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int p = l0 + g;  // element within the chunk
      const bool live = p < e.w;
      Real c = cv[h][g];
      if (decide && live) {
        const Real Sd = S0 * O::ex(mu_d + sg * wd[h][g]);
        if (K - Sd > Real(0)) {
          const Real x = Real(1) + (Sd * invK - Real(1)) * isw_d;
          const Real cont = b0 + x * (b1 + x * (b2 + x * b3));
          const Real ex = (K - Sd) * disc_d;
          if (ex > cont) {
            c = ex;
            cf[p] = c;
          }
        }
      }
      if (k < 1 && live) {  // last pass: the path's final discounted cash flow
        fin[0] += (double)c;
        fin[1] += (double)c * (double)c;
      }
      if (k >= 1) {  // uniform per workgroup
        const Real Wk = ba * wd[h][g] + bb * z[g];
        wk[g] = Wk;
        const Real Sk = S0 * O::ex(mu_k + sg * Wk);
        const bool itm = live && K - Sk > Real(0);
        const Real x = itm ? Real(1) + (Sk * invK - Real(1)) * isw_k : Real(0);
        accum_moments(sm, itm, x, c);
      }
    }
    if (k >= 1) store4(wst + l0, wk);  // padding entries too (never read as paths)
  }
  if (k < 1) {  // last pass of this option: decision only, then its cash-flow sums
    block_sum<2>(fin, lds);
    if (threadIdx.x == 0) {
      a.cfsum[2 * blockIdx.x] = fin[0];
      a.cfsum[2 * blockIdx.x + 1] = fin[1];
    }
    return;
  }
  double s[kNS];
  if constexpr (sizeof(Real) == 4) {
    // FP32 thread sums: FP32 wave trees (6 roundings), FP64 from the wave totals on
    double t[kNM];
    block_sum<kNM>(sm, t, lds);
#pragma unroll
    for (int q = 0; q < kNS; ++q) s[q] = q < kNM ? t[q] : 0.0;
  } else {
#pragma unroll
    for (int q = 0; q < kNS; ++q) s[q] = q < kNM ? (double)sm[q] : 0.0;
    block_sum<kNS>(s, lds);
  }
  if (threadIdx.x == 0) {
    double* dst = a.partials + (size_t)blockIdx.x * kNS;
    for (int q = 0; q < kNS; ++q) dst[q] = s[q];
  }
}

// After pass j: workgroup i takes the i-th option by step count (the options
// still alive at pass j are a prefix), sums its workgroups' partials and
// solves for the coefficients that pass j + 1 applies.
__global__ __launch_bounds__(kBlock) void lsmc_solve_kernel(LsmcArgs a) {
  __shared__ double lds[kNS * (kBlock / 64)];
  const int o = a.order[blockIdx.x];
  const OptMeta m = a.meta[o];
  if (m.steps - a.pass - 1 < 1) return;  // no regression date left (uniform)
  double* bout = a.beta + (size_t)(((a.pass + 1) & 1) * a.n_options + o) * 5;
  double t[kNS];
  combine_partials<kNS>(a.partials, m.first_block, m.nblk, t, lds);
  if (threadIdx.x == 0) lsmc_solve(t, bout);
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void lsmc_final_kernel(LsmcArgs a, double* price2, double* european) {
  __shared__ double lds[2 * (kBlock / 64)];
  const int o = blockIdx.x;
  const OptMeta m = a.meta[o];
  double s[2];
  combine_partials<2>(a.cfsum, m.first_block, m.nblk, s, lds);
  double eu[1] = {0.0};
  if (european) {  // uniform; the barrier orders thread 0's reads of lds before its reuse
    __syncthreads();
    combine_partials<1>(a.eur, m.first_block, m.nblk, eu, lds);
  }
  if (threadIdx.x == 0) {
    const double N = (double)m.paths;
    const double mean = s[0] / N;
    double var = (N * s[1] - s[0] * s[0]) / (N * (N - 1.0));
    const double now = m.K - m.S0 > 0.0 ? m.K - m.S0 : 0.0;
    price2[2 * o] = mean > now ? mean : now;
    price2[2 * o + 1] = sqrt(var > 0.0 ? var : 0.0) / sqrt(N);
    if (european) european[o] = eu[0] / N;
  }
}

__global__ void zero_f64_kernel(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Resident sweep: the whole backward sweep of every option in one cooperative
// launch, path state held in registers
// ---------------------------------------------------------------------------
// An option's paths are cut into parts of up to kThreads * PPT paths; a part
// is one workgroup that keeps its paths' bridge state W and discounted cash
// flow in registers from the terminal date down to date 1, so nothing is
// streamed through HBM between dates.  The workgroup fills a CU (one per CU):
// an option then meets at most 16 partners per date (262,144 paths in FP32),
// and no other option's workgroup shares its CU, whose pace would otherwise
// couple every option on the chip into one lockstep (measured: 256-thread
// parts four to a CU, 22 ms for the 256-put batch; one 1024-thread part per
// CU, 13.7 ms).  The parts of one option meet once per date, and nothing else
// in the grid waits for them:
//   * each part sums its moments (thread sums, wave trees, FP64 across waves)
//     and publishes them in its record for the date's parity as 64-bit words
//     that each carry the date's epoch (pass + 1) in their high half and 32
//     bits of the sums in their low half (FP32: one word per sum; FP64: two,
//     the halves of the double).  A 64-bit store is single-copy atomic, so a
//     word whose epoch matches holds this date's bits and no ordering between
//     the words is needed; the stores are device-scope (sc1, written through
//     to the point of coherence);
//   * wave 0 of every part reads the records of all the option's parts (lane
//     l takes parts l, l + 64, ...) with device-scope loads, all words of a
//     record at once, until every word carries the epoch (one round trip when
//     the part has already published), and the sums are added in part order
//     (every part derives bit-identical coefficients, and reruns reproduce
//     them).  Lane 0 solves by the FP64 Cholesky and broadcasts the
//     coefficients through LDS.
// Two record buffers (pass parity) suffice: a part can publish pass j + 2
// (overwriting pass j) only after every part has published pass j + 1, which
// each does only after it has read all of pass j.
// No agent-scope fences are used: each one writes back or invalidates the
// XCD's L2, and a measured 1024-workgroup barrier cost 20 us per date with
// them against 3 us with device-scope stores and loads of the records.
// Workgroups run host-built task lists (option, part); all parts of an option
// appear in the same global order on every list, and the launch is
// cooperative (all workgroups resident), so every barrier completes.  Each
// poll is bounded: past spin_limit polls (or once another workgroup gave up)
// a workgroup raises the abort flag and returns, so every wave exits; the
// host reports the abort as an error.
// Record words of one part and date: (epoch << 32) | 32 bits of the sums.
template <typename Real>
struct RecWords {
  static constexpr int kN = sizeof(Real) == 4 ? kNM : 2 * kNM;
};
constexpr int kRec = 2 * kNM;  // words per record slot (the FP64 layout; FP32 uses the first kNM)

// Per-option and per-date constants of the resident sweep, computed on the
// host in FP64 (the kernel rounds them to the working precision as the
// streamed sweep does): device FP64 exp() would keep its polynomial
// coefficients in registers beside the paths' state.
struct ResOpt {
  double sqT, muT, discT, invK;  // sqrt(T), (r - sigma^2/2) T, e^{-rT}, 1/K
  int date0, pad;                // first entry of the option's date table (entry d = date d)
};
struct DateConst {
  double mu;    // (r - sigma^2/2) d dt
  double isw;   // 1 / (sigma sqrt(d dt))
  double disc;  // e^{-r d dt}
  double ba;    // bridge d -> d - 1: W_{d-1} = ba W_d + bb Z
  double bb;
};

struct ResArgs {
  const OptMeta* meta;    // first_block = first part, nblk = parts
  const ResOpt* ropt;
  const DateConst* dates;
  const int* slot_begin;  // [grid + 1]: the workgroup's tasks
  const int2* tasks;      // (option, part) in the global option order
  unsigned long long* records;  // [part][2][kRec] tagged words
  double* eur;            // [part] European payoff sums
  double* cfsum;          // [part][2] final cash-flow sum and sum of squares
  int* abort_flag;
  int part_paths;
  int spin_limit;
  uint32_t k0, k1, w;
};

// Wave totals for the resident kernel, in a fixed order (every lane ends with
// the same bits): four DPP steps inside each row of 16 lanes, a swizzle
// across the 32-lane halves and the two halves' totals read as scalars.
// Unlike a ds_bpermute tree they need no per-lane address registers, which
// the compiler would otherwise hoist out of the sweep's loops and keep live
// beside the paths' state.
template <int kCtrl>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, true);
}
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = dpp_i32<kCtrl>((int)b), hi = dpp_i32<kCtrl>((int)(b >> 32));
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double swz16_f64(double v) {  // lane ^ 16 within each 32-lane half
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_ds_swizzle((int)b, 0x401F), hi = __builtin_amdgcn_ds_swizzle((int)(b >> 32), 0x401F);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_f64<0xB1>(v);   // lane ^ 1
  v += dpp_f64<0x4E>(v);   // lane ^ 2
  v += dpp_f64<0x141>(v);  // row half mirror
  v += dpp_f64<0x140>(v);  // row mirror
  v += swz16_f64(v);
  return readlane_f64(v, 0) + readlane_f64(v, 32);
}
__device__ __forceinline__ float wave_total(float v) {
  v += dpp_f32<0xB1>(v);
  v += dpp_f32<0x4E>(v);
  v += dpp_f32<0x141>(v);
  v += dpp_f32<0x140>(v);
  v += __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x401F));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0)) +
         __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
}

// Workgroup totals of NV values per thread: wave totals (working precision),
// combined over the kThreads / 64 waves in FP64 in wave order by lane q of
// wave 0 for value q.  The total of value q is returned in thread q < NV
// (other threads: unspecified); the caller separates reuses of lds by a
// barrier.
// Wave totals of up to 12 FP32 values per lane, transposed: v_permlane32_swap
// pairs value j with value j + 6 (each half-wave then sums one of them over
// both halves), v_permlane16_swap pairs j with j + 3 within the halves, and
// four DPP steps finish each of the 3 remaining values within its row of 16
// lanes.  Row r (lanes 16 r .. 16 r + 15) ends with the totals of values
// 3 r + j in t[j], 35 instructions where one total per value costs ~9.
template <int NV>
__device__ __forceinline__ void wave_totals_transposed(const float (&v)[NV], float (&t)[3]) {
  static_assert(NV <= 12, "four rows of three values");
  auto bits = [](float x) { return __builtin_bit_cast(unsigned, x); };
  auto flt = [](unsigned x) { return __builtin_bit_cast(float, x); };
  float u[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const auto r = __builtin_amdgcn_permlane32_swap(bits(v[j]), bits(j + 6 < NV ? v[j + 6] : 0.f), false, false);
    u[j] = flt(r[0]) + flt(r[1]);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const auto r = __builtin_amdgcn_permlane16_swap(bits(u[j]), bits(u[j + 3]), false, false);
    t[j] = flt(r[0]) + flt(r[1]);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    t[j] += dpp_f32<0xB1>(t[j]);   // lane ^ 1
    t[j] += dpp_f32<0x4E>(t[j]);   // lane ^ 2
    t[j] += dpp_f32<0x141>(t[j]);  // row half mirror
    t[j] += dpp_f32<0x140>(t[j]);  // row mirror
  }
}

template <int kThreads, typename T, int NV>
__device__ __forceinline__ double res_block_sum(const T (&v)[NV], double* lds) {
  constexpr int kWaves = kThreads / 64;
  const int wave = threadIdx.x >> 6;
  if constexpr (sizeof(T) == 4 && NV > 4) {
    float t[3];
    wave_totals_transposed(v, t);
    const int lane = threadIdx.x & 63;
    if ((lane & 15) == 0) {  // row r = lane / 16 holds values 3 r .. 3 r + 2
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const int q = 3 * (lane >> 4) + j;
        if (q < NV) lds[q * kWaves + wave] = (double)t[j];
      }
    }
  } else {
    double w[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) w[q] = (double)wave_total(v[q]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int q = 0; q < NV; ++q) lds[q * kWaves + wave] = w[q];
    }
  }
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x < NV) {
#pragma unroll
    for (int i = 0; i < kWaves; ++i) t += lds[threadIdx.x * kWaves + i];
  }
  return t;
}

template <typename Real, int PPT, int kThreads>
__global__ __launch_bounds__(kThreads) void lsmc_resident_kernel(ResArgs a) {
  using O = MathOps<Real>;
  constexpr int NG = PPT / 4;
  static_assert(64 + 64 * NG <= kThreads, "waves 1 .. NG take wave 0's normals");
  __shared__ double lds[kNS * (kThreads / 64)];
  __shared__ double stage[64 * kNM];  // records of up to 64 parts, [part][moment]
  __shared__ double sums[kNS];
  __shared__ double beta_lds[5];
  __shared__ int stop_lds;
  // a window of the option's date constants, dates [wbase, wbase +
  // kDatesLds): a load from device memory per date would put its latency on
  // every date's critical path.  Options with more dates slide the window
  // down when the sweep leaves it.
  constexpr int kDatesLds = 512;
  __shared__ DateConst dl[kDatesLds];
  // thread-private slots (no barrier orders them): the next bridge step's
  // normals of each of this thread's path groups, generated while the part
  // waits for its partners' records, off the date's critical path; and the
  // spots of the date awaiting its decision, made by the bridge step (so the
  // decision and the regression see the same bits) and kept here rather than
  // in registers beside the bridge states and cash flows
  typedef typename Quad<Real>::V V4;
  __shared__ V4 znext[NG][kThreads];
  __shared__ V4 spot[NG][kThreads];
  const int lane = threadIdx.x & 63;
  // Group h of this thread holds paths 4 (h kThreads + threadIdx.x) + g of
  // the part: its Philox counter word is ctr0 + h kThreads.  The base is
  // passed through an empty asm at each use, so the optimiser recomputes the
  // groups' counters there instead of keeping NG of them live in the loops.
  auto gen_normals = [&](int idx, int ctr0, int o) __attribute__((always_inline)) {  // Philox step index idx
#pragma unroll
    for (int h = 0; h < NG; ++h) {
      int c = ctr0;
      __asm__ volatile("" : "+v"(c));
      Real z[4];
      normals4<Real>(U4{(uint32_t)(c + h * kThreads), (uint32_t)idx, (uint32_t)o, a.w}, a.k0, a.k1, z[0], z[1], z[2],
                     z[3]);
      V4 v;
      v[0] = z[0];
      v[1] = z[1];
      v[2] = z[2];
      v[3] = z[3];
      znext[h][threadIdx.x] = v;
    }
  };
  const int t_end = a.slot_begin[blockIdx.x + 1];
  for (int ti = a.slot_begin[blockIdx.x]; ti < t_end; ++ti) {
    const int2 task = a.tasks[ti];
    const int o = task.x, part = task.y;
    const OptMeta m = a.meta[o];
    const int first = part * a.part_paths;
    const int np = min(a.part_paths, m.paths - first);
    // path 4 (h kThreads + threadIdx.x) + g is live iff 4 h kThreads + g < rem
    const int rem = np - 4 * (int)threadIdx.x;
    const int ctr0 = (first >> 2) + (int)threadIdx.x;  // part_paths is a multiple of 4
    const int rec = m.first_block + part;
    const ResOpt ro = a.ropt[o];
    int wbase = max(0, m.steps + 1 - kDatesLds);  // uniform
    auto load_window = [&]() {
      for (int i = threadIdx.x; i < kDatesLds && wbase + i <= m.steps; i += kThreads) dl[i] = a.dates[ro.date0 + wbase + i];
    };
    load_window();
    // (the barrier in the terminal sum below orders the copy before the dates' reads)
    auto date = [&](int i) -> DateConst { return dl[i - wbase]; };
    const Real K = (Real)m.K, invK = (Real)ro.invK, S0 = (Real)m.S0, sg = (Real)(m.sigma * O::kE);
    // per path: bridge state W and discounted cash flow C (the spot of the
    // date awaiting its decision is in spot[][])
    Real W[NG][4], C[NG][4];
    {
      const Real sqT = (Real)ro.sqT, muT = (Real)(ro.muT * O::kE), discT = (Real)ro.discT;
      double eu[1] = {0.0};
#pragma unroll
      for (int h = 0; h < NG; ++h) {
        Real z[4];  // computed for padding groups too (never used for them): no branch per group
        normals4<Real>(U4{(uint32_t)(ctr0 + h * kThreads), (uint32_t)(m.steps - 1), (uint32_t)o, a.w}, a.k0, a.k1,
                       z[0], z[1], z[2], z[3]);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const bool live = 4 * h * kThreads + g < rem;
          W[h][g] = live ? sqT * z[g] : Real(0);
          const Real ST = S0 * O::ex(muT + sg * W[h][g]);
          C[h][g] = live ? discT * (K - ST > Real(0) ? K - ST : Real(0)) : Real(0);
          eu[0] += (double)C[h][g];
        }
        __builtin_amdgcn_sched_barrier(0);  // one group of paths at a time: bounds the live registers
      }
      const double t = res_block_sum<kThreads, double, 1>(eu, lds);
      if (threadIdx.x == 0) a.eur[rec] = t;
      __syncthreads();  // lds is reused by the first pass
      if (m.steps >= 2) gen_normals(m.steps - 2, ctr0, o);  // the first bridge step's (N -> N - 1)
    }
    for (int pass = 0; pass < m.steps; ++pass) {
      const int d = m.steps - pass;  // decision date
      const int k = d - 1;           // regression date
      if (k < wbase) {  // uniform: the window slides down to end at date d
        // (every read of the previous window ended before the last pass's barrier)
        wbase = max(0, d + 1 - kDatesLds);
        load_window();
        __syncthreads();
      }
      if (pass >= 1 && beta_lds[4] != 0.0) {  // exercise decisions of date d (uniform; per path by selects)
        const Real b0 = (Real)beta_lds[0], b1 = (Real)beta_lds[1], b2 = (Real)beta_lds[2], b3 = (Real)beta_lds[3];
        const DateConst cd = date(d);
        const Real isw_d = (Real)cd.isw, disc_d = (Real)cd.disc;
#pragma unroll
        for (int h = 0; h < NG; ++h) {
          const V4 sv = spot[h][threadIdx.x];
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const Real Sd = sv[g];
            const Real x = Real(1) + (Sd * invK - Real(1)) * isw_d;
            const Real cont = b0 + x * (b1 + x * (b2 + x * b3));
            const Real ex = (K - Sd) * disc_d;
            // padding paths never exercise
            const bool live = 4 * h * kThreads + g < rem;
            C[h][g] = live && K - Sd > Real(0) && ex > cont ? ex : C[h][g];
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      if (k < 1) {  // decision at date 1 done: the part's final cash-flow sums
        double fin[2] = {0.0, 0.0};
#pragma unroll
        for (int h = 0; h < NG; ++h) {
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            fin[0] += (double)C[h][g];  // padding paths hold 0
            fin[1] += (double)C[h][g] * (double)C[h][g];
          }
        }
        const double t = res_block_sum<kThreads, double, 2>(fin, lds);
        if (threadIdx.x < 2) a.cfsum[2 * rec + threadIdx.x] = t;
        __syncthreads();  // lds is reused by the next task
        break;
      }
      // bridge to date k and the regression moments of date k
      Real sm[kNM];
#pragma unroll
      for (int q = 0; q < kNM; ++q) sm[q] = Real(0);
      {
        const DateConst cd = date(d), ck = date(k);
        const Real mu_k = (Real)(ck.mu * O::kE), isw_k = (Real)ck.isw;
        const Real ba = (Real)cd.ba, bb = (Real)cd.bb;
#pragma unroll
        for (int h = 0; h < NG; ++h) {
          const V4 zv = znext[h][threadIdx.x];  // Philox step index k - 1, generated last date
          const Real z[4] = {zv[0], zv[1], zv[2], zv[3]};
          V4 sk;
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const bool live = 4 * h * kThreads + g < rem;
            const Real Wk = ba * W[h][g] + bb * z[g];
            W[h][g] = live ? Wk : Real(0);
            const Real Sk = S0 * O::ex(mu_k + sg * Wk);
            sk[g] = Sk;
            const bool itm = live && K - Sk > Real(0);
            const Real x = itm ? Real(1) + (Sk * invK - Real(1)) * isw_k : Real(0);
            accum_moments(sm, itm, x, C[h][g]);
          }
          spot[h][threadIdx.x] = sk;
          __builtin_amdgcn_sched_barrier(0);
        }
      }
      // publish this part's moments, generate the next bridge step's normals,
      // then gather every part's moments
      constexpr int NW = RecWords<Real>::kN;
      const unsigned long long tag = (unsigned long long)(pass + 1) << 32;
      {
        const double s = res_block_sum<kThreads, Real, kNM>(sm, lds);
        const int q = threadIdx.x;  // lane q < kNM publishes moment q
        if (q < kNM) {
          unsigned long long* r = a.records + ((size_t)rec * 2 + (pass & 1)) * kRec;
          if constexpr (sizeof(Real) == 4) {
            __hip_atomic_store(r + q, tag | __builtin_bit_cast(unsigned, (float)s), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
          } else {
            const unsigned long long b = __builtin_bit_cast(unsigned long long, s);
            __hip_atomic_store(r + 2 * q, tag | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(r + 2 * q + 1, tag | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          }
        }
      }
      // the bridge k -> k - 1 of the next date.  FP32: wave 0 goes straight
      // to the poll and the solve below, the other waves generate its normals
      // as well as their own (waves 1 .. NG one (thread, group) slot of wave 0
      // per thread), so the generation stays off the date's critical path.
      // FP64 generation (software log and sincos) outlasts the poll and the
      // solve: there the extra share would lengthen the date, and every wave
      // generates its own.
      constexpr bool kOffload = sizeof(Real) == 4;
      if (k >= 2 && (!kOffload || threadIdx.x >= 64)) {  // uniform per wave
        gen_normals(k - 2, ctr0, o);
        const int j = (int)threadIdx.x - 64;
        if (kOffload && j < 64 * NG) {
          const int h = j >> 6, t = j & 63;
          Real z[4];
          normals4<Real>(U4{(uint32_t)((first >> 2) + t + h * kThreads), (uint32_t)(k - 2), (uint32_t)o, a.w}, a.k0,
                         a.k1, z[0], z[1], z[2], z[3]);
          V4 v;
          v[0] = z[0];
          v[1] = z[1];
          v[2] = z[2];
          v[3] = z[3];
          znext[h][t] = v;
        }
      }
      if (threadIdx.x < 64) {
        // lanes l < parts poll and stage part l's record (rounds of 64 parts);
        // lanes q < kNM then add moment q over the staged parts in part order
        double acc = 0.0;
        int bad = 0;
        for (int base = 0; base < m.nblk && !bad; base += 64) {
          const int p = base + lane;
          if (p < m.nblk) {
            unsigned long long* r = a.records + ((size_t)(m.first_block + p) * 2 + (pass & 1)) * kRec;
            unsigned long long w[NW];
            // read all words of the record, until every one carries the epoch
            for (int spins = 0;; ++spins) {
#pragma unroll
              for (int q = 0; q < NW; ++q) w[q] = __hip_atomic_load(r + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              bool all = true;
#pragma unroll
              for (int q = 0; q < NW; ++q) all = all && (w[q] & 0xffffffff00000000ull) == tag;
              if (all) break;
              // the abort flag is read once per 64 polls
              if ((spins & 63) == 63 &&
                  (spins > a.spin_limit || __hip_atomic_load(a.abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(a.abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                bad = 1;
                break;
              }
              __builtin_amdgcn_s_sleep(2);
            }
            if (!bad) {
#pragma unroll
              for (int q = 0; q < kNM; ++q) {
                if constexpr (sizeof(Real) == 4)
                  stage[lane * kNM + q] = (double)__builtin_bit_cast(float, (unsigned)w[q]);
                else
                  stage[lane * kNM + q] =
                      __builtin_bit_cast(double, (w[2 * q] & 0xffffffffull) | (w[2 * q + 1] << 32));
              }
            }
          }
          bad = __any(bad);
          __builtin_amdgcn_wave_barrier();  // the wave's LDS writes land before its reads
          if (!bad && lane < kNM) {
            const int n = min(64, m.nblk - base);
            for (int i = 0; i < n; ++i) acc += stage[i * kNM + lane];
          }
          __builtin_amdgcn_wave_barrier();  // reads done before the next round's writes
        }
        if (lane < kNM) sums[lane] = acc;
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
          if (!bad) {
            double t[kNS];
#pragma unroll
            for (int q = 0; q < kNS; ++q) t[q] = q < kNM ? sums[q] : 0.0;
            lsmc_solve(t, beta_lds);
          }
          stop_lds = bad;
        }
      }
      __syncthreads();
      if (stop_lds) return;  // uniform: every thread read the same flag after the barrier
    }
  }
}

__global__ void lsmc_resident_clear_kernel(unsigned long long* records, size_t n, int* abort_flag) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) records[i] = 0ull;  // epoch 0: no date published
  if (i == 0) *abort_flag = 0;
}

// ---------------------------------------------------------------------------
// Reference algorithm, one option
// ---------------------------------------------------------------------------
struct RefArgs {
  int n_paths, n_steps;
  double spot, strike;
  double fwd_l2e, vsdt_l2e;  // main.cpp:55-59, log2 domain as perform_sims
  const double* z;           // buffer / MRG32k3a modes: step s-1 of path p at z[(s-1) * zrow + p]
  long long zrow;            // n_paths (caller buffer) or n_paths / 1024 (the sample's overlapping streams)
  uint32_t k0, k1, w;
  void* a;                   // Real [n_steps + 1][n_paths]
  void* cf;                  // Real [n_paths]
  double* partials;          // [workgroup][8]
  int n_partials;            // workgroups of the step kernel
  double* beta;              // [2][3] double-buffered by step parity
  double* out;               // [2]: american, european
};

template <typename Real, bool kBuffer>
__global__ __launch_bounds__(kBlock) void ref_paths_kernel(RefArgs r) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= r.n_paths) return;
  Real* a = (Real*)r.a;
  const size_t N = (size_t)r.n_paths;
  Real x = (Real)r.spot;
  a[p] = x;
  Real z4[4] = {0, 0, 0, 0};
  for (int s = 1; s <= r.n_steps; ++s) {
    Real z;
    if (kBuffer) {
      z = (Real)r.z[(size_t)(s - 1) * (size_t)r.zrow + p];
    } else {
      const int j = s - 1;
      if ((j & 3) == 0)
        normals4<Real>(U4{(uint32_t)p, (uint32_t)(j >> 2), 0u, r.w}, r.k0, r.k1, z4[0], z4[1], z4[2], z4[3]);
      z = pick4(z4, j & 3);
    }
    // utilities.cpp:83: a[s] = a[s-1] * exp2(fwdFactor + volSqrtdt * z)
    x = x * fexp2((Real)r.fwd_l2e + (Real)r.vsdt_l2e * z);
    a[(size_t)s * N + p] = x;
  }
}

// Step s of main.cpp:131-266: decision with the previous step's betas (or
// the 1/N-scaled terminal cash flow at s = n_steps - 1), then the eight ITM
// sums; the last workgroup inverts the ridge-corrected 3x3 moment matrix.
template <typename Real>
__global__ __launch_bounds__(kBlock) void ref_step_kernel(RefArgs r, int s) {
  __shared__ double lds[8 * (kBlock / 64)];
  const size_t N = (size_t)r.n_paths;
  const Real* ap = (const Real*)r.a + (size_t)s * N;
  Real* cf = (Real*)r.cf;
  const double* bin = r.beta + (s & 1) * 3;
  const Real b0 = (Real)bin[0], b1 = (Real)bin[1], b2 = (Real)bin[2];
  const Real K = (Real)r.strike;
  const Real avg = (Real)(1.0 / r.n_paths);
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x; p < N; p += (size_t)gridDim.x * kBlock) {
    const Real x = ap[p];
    Real c = cf[p];
    if (s < r.n_steps - 1) {
      if (x <= K) {
        const Real reg = b0 + b1 * x + b2 * (x * x);
        const Real payoff = K - x;
        if (payoff > reg) c = payoff;
      }
    } else {
      c = (K - x > Real(0) ? K - x : Real(0)) * avg;
    }
    cf[p] = c;
    if (x <= K) {
      const double xd = (double)x, cd = (double)c, x2 = xd * xd;
      v[0] += 1.0;
      v[1] += xd;
      v[2] += x2;
      v[3] += x2 * xd;
      v[4] += x2 * x2;
      v[5] += cd;
      v[6] += cd * xd;
      v[7] += cd * x2;
    }
  }
  block_sum<8>(v, lds);
  if (threadIdx.x == 0)
    for (int q = 0; q < 8; ++q) r.partials[(size_t)blockIdx.x * 8 + q] = v[q];
}

// One workgroup after each step kernel: the regression of step s.
__global__ __launch_bounds__(kBlock) void ref_solve_kernel(RefArgs r, int s) {
  __shared__ double lds[8 * (kBlock / 64)];
  double* bout = r.beta + ((s + 1) & 1) * 3;
  double t[8];
  combine_partials<8>(r.partials, 0, r.n_partials, t, lds);
  if (threadIdx.x == 0) {
    const double ridge = 0.01;  // main.cpp:244-246
    const double s0 = t[0] + ridge, s1 = t[1] + ridge, s2 = t[2] + ridge, s3 = t[3] + ridge, s4 = t[4] + ridge;
    const double x[3][3] = {{s0, s1, s2}, {s1, s2, s3}, {s2, s3, s4}};
    // utilities.cpp:90-123 cofactor inverse
    const double det = x[0][0] * x[1][1] * x[2][2] + x[1][0] * x[2][1] * x[0][2] + x[2][0] * x[0][1] * x[1][2] -
                       x[2][0] * x[1][1] * x[0][2] - x[1][0] * x[0][1] * x[2][2] - x[0][0] * x[2][1] * x[1][2];
    double inv[3][3];
    inv[0][0] = (x[1][1] * x[2][2] - x[2][1] * x[1][2]) / det;
    inv[0][1] = (x[0][2] * x[2][1] - x[0][1] * x[2][2]) / det;
    inv[0][2] = (x[0][1] * x[1][2] - x[0][2] * x[1][1]) / det;
    inv[1][0] = (x[1][2] * x[2][0] - x[1][0] * x[2][2]) / det;
    inv[1][1] = (x[0][0] * x[2][2] - x[0][2] * x[2][0]) / det;
    inv[1][2] = (x[0][2] * x[1][0] - x[0][0] * x[1][2]) / det;
    inv[2][0] = (x[1][0] * x[2][1] - x[1][1] * x[2][0]) / det;
    inv[2][1] = (x[0][1] * x[2][0] - x[0][0] * x[2][1]) / det;
    inv[2][2] = (x[0][0] * x[1][1] - x[0][1] * x[1][0]) / det;
    // main.cpp:259-261
    bout[0] = t[5] * inv[0][0] + t[6] * inv[1][0] + t[7] * inv[2][0];
    bout[1] = t[5] * inv[0][1] + t[6] * inv[1][1] + t[7] * inv[2][1];
    bout[2] = t[5] * inv[0][2] + t[6] * inv[1][2] + t[7] * inv[2][2];
  }
}

// main.cpp:269-279: American = sum(cf)/N, European = sum((K - S_T)^+)/N.
template <typename Real>
__global__ __launch_bounds__(kBlock) void ref_final_kernel(RefArgs r) {
  __shared__ double lds[2 * (kBlock / 64)];
  const size_t N = (size_t)r.n_paths;
  const Real* aT = (const Real*)r.a + (size_t)r.n_steps * N;
  const Real* cf = (const Real*)r.cf;
  double v[2] = {0.0, 0.0};
  for (size_t p = threadIdx.x; p < N; p += kBlock) {
    v[0] += (double)cf[p];
    const double e = r.strike - (double)aT[p];
    v[1] += e > 0.0 ? e : 0.0;
  }
  block_sum<2>(v, lds);
  if (threadIdx.x == 0) {
    r.out[0] = v[0] / r.n_paths;
    r.out[1] = v[1] / r.n_paths;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct LsmcPlanImpl {
  int n;
  int precision;
  fsmc_mode mode;
  int sweep = FSMC_LSMC_SWEEP_STREAMED;
  std::vector<OptMeta> meta;
  std::vector<int4> table;
  std::vector<int> order;         // option indices sorted by steps (descending)
  std::vector<int> grid_of_pass;  // workgroups launched by backward pass j
  std::vector<int> live_of_pass;  // options with a date left at backward pass j
  int max_steps = 0;
  size_t real = 4, cf_elems = 0;
  size_t off_meta, off_table, off_order, off_w, off_cf, off_part, off_beta, off_eur, off_cfsum, off_out, ws = 0;
  // resident sweep: meta (first_block / nblk = parts), task lists per workgroup
  int grid = 0, part_paths = 0, parts = 0;
  std::vector<int> slot_begin;
  std::vector<int2> tasks;
  std::vector<ResOpt> ropt;
  std::vector<DateConst> dates;
  size_t off_slots = 0, off_tasks = 0, off_rec = 0, off_abort = 0, off_ropt = 0, off_dates = 0;
};

// Shape of the resident sweep: one workgroup per CU, as wide as the register
// file allows, so an option meets few partners per date and no other
// option's workgroup shares (and paces) its CU.  FP32 keeps 16 paths' bridge
// state and cash flow (32 VGPRs) per thread at 1024 threads (16 waves, four
// per SIMD: <= 128 VGPRs, no spills); FP64 keeps 16 paths (64 VGPRs) at 512
// threads (two waves per SIMD).
constexpr int kResPPT32 = 16, kResThreads32 = 1024;
constexpr int kResPPT64 = 16, kResThreads64 = 512;
constexpr int kResSpinLimit = 1 << 22;  // polls of one record (each ~1 us) before a workgroup gives up

template <typename Real>
const void* resident_kernel() {
  return sizeof(Real) == 4 ? (const void*)lsmc_resident_kernel<float, kResPPT32, kResThreads32>
                           : (const void*)lsmc_resident_kernel<double, kResPPT64, kResThreads64>;
}
inline int resident_threads(int precision) { return precision == FSMC_FP64 ? kResThreads64 : kResThreads32;
}

// Workgroups of the resident grid on ctx's device (0 when a cooperative
// launch is not available).
int resident_grid(int device, int precision) {
  int coop = 0, cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess || !coop) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
  const void* f = precision == FSMC_FP64 ? resident_kernel<double>() : resident_kernel<float>();
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, resident_threads(precision), 0) != hipSuccess) return 0;
  return cus * per_cu;
}

// Resident sweep plan: parts of part_paths paths; longest-first list
// schedule of the options over the grid's workgroups (an option takes the
// nblk workgroups that free up first and holds them for its steps dates).
// Every list holds its tasks in the global option order.  Returns false if
// an option has more parts than the grid has workgroups.
bool lsmc_make_resident(LsmcPlanImpl* pl, int grid) {
  const int n = pl->n;
  pl->part_paths = pl->precision == FSMC_FP64 ? kResThreads64 * kResPPT64 : kResThreads32 * kResPPT32;
  pl->grid = grid;
  int parts = 0;
  for (int idx = 0; idx < n; ++idx) {
    OptMeta& q = pl->meta[pl->order[idx]];
    q.first_block = parts;
    q.nblk = (q.paths + pl->part_paths - 1) / pl->part_paths;
    if (q.nblk > grid) return false;
    parts += q.nblk;
  }
  pl->parts = parts;
  std::vector<double> avail(grid, 0.0);
  std::vector<std::vector<int2>> lists(grid);
  std::vector<int> slot(grid);
  for (int idx = 0; idx < n; ++idx) {
    const int o = pl->order[idx];
    const OptMeta& q = pl->meta[o];
    std::iota(slot.begin(), slot.end(), 0);
    std::partial_sort(slot.begin(), slot.begin() + q.nblk, slot.end(),
                      [&](int x, int y) { return avail[x] < avail[y] || (avail[x] == avail[y] && x < y); });
    double start = 0.0;
    for (int j = 0; j < q.nblk; ++j) start = std::max(start, avail[slot[j]]);
    for (int j = 0; j < q.nblk; ++j) {
      avail[slot[j]] = start + q.steps;
      lists[slot[j]].push_back(make_int2(o, j));
    }
  }
  pl->slot_begin.assign(grid + 1, 0);
  pl->tasks.clear();
  for (int g = 0; g < grid; ++g) {
    pl->tasks.insert(pl->tasks.end(), lists[g].begin(), lists[g].end());
    pl->slot_begin[g + 1] = (int)pl->tasks.size();
  }
  // the same FP64 expressions as the streamed sweep's kernels
  pl->ropt.resize(n);
  pl->dates.clear();
  for (int o = 0; o < n; ++o) {
    const OptMeta& q = pl->meta[o];
    ResOpt& ro = pl->ropt[o];
    ro.sqT = sqrt(q.T);
    ro.muT = (q.r - 0.5 * q.sigma * q.sigma) * q.T;
    ro.discT = exp(-q.r * q.T);
    ro.invK = 1.0 / q.K;
    ro.date0 = (int)pl->dates.size();
    ro.pad = 0;
    const double dt = q.T / q.steps, mu = (q.r - 0.5 * q.sigma * q.sigma) * dt;
    for (int d = 0; d <= q.steps; ++d) {
      DateConst c;
      c.mu = mu * d;
      c.isw = d >= 1 ? 1.0 / (q.sigma * sqrt(d * dt)) : 0.0;
      c.disc = exp(-q.r * d * dt);
      c.ba = d >= 1 ? (double)(d - 1) / d : 0.0;
      c.bb = d >= 1 ? sqrt(dt * (d - 1) / d) : 0.0;
      pl->dates.push_back(c);
    }
  }
  return true;
}

int lsmc_make_plan(int n, const fsmc_mode* m, const double* S0, const double* K, const double* sigma,
                   const double* r, const double* T, const int* steps, const int* paths, int sweep, int grid,
                   LsmcPlanImpl* pl) {
  FSMC_REQUIRE(n >= 1 && n <= (1 << 20), "n_options %d outside [1, 2^20]", n);
  FSMC_REQUIRE(sweep == FSMC_LSMC_SWEEP_AUTO || sweep == FSMC_LSMC_SWEEP_RESIDENT || sweep == FSMC_LSMC_SWEEP_STREAMED,
               "sweep %d is not an FSMC_LSMC_SWEEP_* value", sweep);
  FSMC_TRY(validate_mode(m, false));
  FSMC_REQUIRE(S0 && K && sigma && r && T && steps && paths, "NULL host pointer");
  for (int o = 0; o < n; ++o) {
    FSMC_REQUIRE(steps[o] >= 1 && paths[o] >= 2 && paths[o] <= (1 << 30),
                 "option %d: steps %d / paths %d out of range", o, steps[o], paths[o]);
    FSMC_REQUIRE(S0[o] > 0 && K[o] > 0 && sigma[o] > 0 && T[o] > 0, "option %d: S0, K, sigma, T must be > 0", o);
  }
  pl->n = n;
  pl->mode = *m;
  pl->precision = m->precision;
  pl->real = m->precision == FSMC_FP64 ? 8 : 4;
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return steps[x] > steps[y]; });
  pl->meta.resize(n);
  pl->table.clear();
  long long cfo = 0;
  for (int idx = 0; idx < n; ++idx) {
    const int o = order[idx];
    OptMeta& q = pl->meta[o];
    q.cf_off = cfo;
    q.steps = steps[o];
    q.paths = paths[o];
    q.nblk = (paths[o] + kChunk - 1) / kChunk;
    q.first_block = (int)pl->table.size();
    q.S0 = S0[o];
    q.K = K[o];
    q.sigma = sigma[o];
    q.r = r[o];
    q.T = T[o];
    for (int c = 0; c < q.nblk; ++c)
      pl->table.push_back(make_int4(o, c, (int)(q.cf_off + (long long)c * kChunk), std::min(kChunk, paths[o] - c * kChunk)));
    cfo += (paths[o] + 3) / 4 * 4;  // segments aligned and padded to whole groups of 4
  }
  // table entries address cash flows with int offsets
  FSMC_REQUIRE(cfo <= (long long)0x7fffffff, "total paths %lld over 2^31", cfo);
  pl->cf_elems = (size_t)cfo;
  pl->max_steps = steps[order[0]];
  pl->order = order;
  pl->grid_of_pass.assign(pl->max_steps, 0);
  pl->live_of_pass.assign(pl->max_steps, 0);
  // options sorted by steps: those with steps > j are a prefix of the table
  for (int j = 0; j < pl->max_steps; ++j) {
    int blocks = 0, live = 0;
    for (int idx = 0; idx < n && steps[order[idx]] > j; ++idx) {
      blocks += pl->meta[order[idx]].nblk;
      ++live;
    }
    pl->grid_of_pass[j] = blocks;
    pl->live_of_pass[j] = live;
  }
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes, 256);
    return at;
  };
  pl->sweep = FSMC_LSMC_SWEEP_STREAMED;
  if (sweep != FSMC_LSMC_SWEEP_STREAMED) {
    FSMC_REQUIRE(grid > 0 || sweep == FSMC_LSMC_SWEEP_AUTO, "the resident sweep needs a cooperative launch");
    const std::vector<OptMeta> streamed = pl->meta;
    if (grid > 0 && lsmc_make_resident(pl, grid)) {
      pl->sweep = FSMC_LSMC_SWEEP_RESIDENT;
      pl->off_meta = take(sizeof(OptMeta) * n);
      pl->off_slots = take(sizeof(int) * (grid + 1));
      pl->off_tasks = take(sizeof(int2) * pl->tasks.size());
      pl->off_rec = take(sizeof(unsigned long long) * 2 * kRec * pl->parts);
      pl->off_eur = take(sizeof(double) * pl->parts);
      pl->off_cfsum = take(sizeof(double) * 2 * pl->parts);
      pl->off_abort = take(sizeof(int));
      pl->off_ropt = take(sizeof(ResOpt) * n);
      pl->off_dates = take(sizeof(DateConst) * pl->dates.size());
      pl->off_out = take(sizeof(double) * 3 * n);
      pl->ws = off;
      return FSMC_OK;
    }
    FSMC_REQUIRE(sweep == FSMC_LSMC_SWEEP_AUTO, "an option needs more than the %d resident workgroups", grid);
    pl->meta = streamed;
  }
  pl->off_meta = take(sizeof(OptMeta) * n);
  pl->off_table = take(sizeof(int4) * pl->table.size());
  pl->off_order = take(sizeof(int) * n);
  pl->off_w = take(pl->real * pl->cf_elems);
  pl->off_cf = take(pl->real * pl->cf_elems);
  pl->off_part = take(sizeof(double) * kNS * pl->table.size());
  pl->off_beta = take(sizeof(double) * 10 * n);
  pl->off_eur = take(sizeof(double) * pl->table.size());
  pl->off_cfsum = take(sizeof(double) * 2 * pl->table.size());
  pl->off_out = take(sizeof(double) * 3 * n);
  pl->ws = off;
  return FSMC_OK;
}

template <typename Real>
int lsmc_enqueue_resident(const LsmcPlanImpl& pl, char* ws, hipStream_t s) {
  ResArgs r;
  r.meta = (const OptMeta*)(ws + pl.off_meta);
  r.ropt = (const ResOpt*)(ws + pl.off_ropt);
  r.dates = (const DateConst*)(ws + pl.off_dates);
  r.slot_begin = (const int*)(ws + pl.off_slots);
  r.tasks = (const int2*)(ws + pl.off_tasks);
  r.records = (unsigned long long*)(ws + pl.off_rec);
  r.eur = (double*)(ws + pl.off_eur);
  r.cfsum = (double*)(ws + pl.off_cfsum);
  r.abort_flag = (int*)(ws + pl.off_abort);
  r.part_paths = pl.part_paths;
  r.spin_limit = kResSpinLimit;
  r.k0 = (uint32_t)pl.mode.seed;
  r.k1 = (uint32_t)(pl.mode.seed >> 32);
  r.w = counter_w(kTagLsmc, pl.mode.stream);
  const size_t nrec = (size_t)2 * kRec * pl.parts;
  hipLaunchKernelGGL(lsmc_resident_clear_kernel, dim3((unsigned)((nrec + 255) / 256)), dim3(256), 0, s, r.records,
                     nrec, r.abort_flag);
  void* args[] = {&r};
  FSMC_HIP(hipLaunchCooperativeKernel(resident_kernel<Real>(), dim3(pl.grid), dim3(resident_threads(pl.precision)),
                                      args, 0, s));
  LsmcArgs a{};
  a.meta = r.meta;
  a.eur = r.eur;
  a.cfsum = r.cfsum;
  double* out = (double*)(ws + pl.off_out);
  hipLaunchKernelGGL(lsmc_final_kernel<Real>, dim3(pl.n), dim3(kBlock), 0, s, a, out, out + 2 * pl.n);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

template <typename Real>
int lsmc_enqueue(const LsmcPlanImpl& pl, char* ws, hipStream_t s) {
  if (pl.sweep == FSMC_LSMC_SWEEP_RESIDENT) return lsmc_enqueue_resident<Real>(pl, ws, s);
  LsmcArgs a;
  a.meta = (const OptMeta*)(ws + pl.off_meta);
  a.table = (const int4*)(ws + pl.off_table);
  a.state = ws + pl.off_w;
  a.cf = ws + pl.off_cf;
  a.partials = (double*)(ws + pl.off_part);
  a.order = (const int*)(ws + pl.off_order);
  a.beta = (double*)(ws + pl.off_beta);
  a.eur = (double*)(ws + pl.off_eur);
  a.cfsum = (double*)(ws + pl.off_cfsum);
  a.n_options = pl.n;
  a.k0 = (uint32_t)pl.mode.seed;
  a.k1 = (uint32_t)(pl.mode.seed >> 32);
  a.w = counter_w(kTagLsmc, pl.mode.stream);
  a.pass = 0;
  hipLaunchKernelGGL(zero_f64_kernel, dim3((10 * pl.n + 255) / 256), dim3(256), 0, s, a.beta, 10 * pl.n);
  const int nb = (int)pl.table.size();
  hipLaunchKernelGGL(lsmc_init_kernel<Real>, dim3(nb), dim3(kBlock), 0, s, a);
  FSMC_HIP(hipGetLastError());
  for (int j = 0; j < pl.max_steps; ++j) {
    a.pass = j;
    hipLaunchKernelGGL(lsmc_pass_kernel<Real>, dim3(pl.grid_of_pass[j]), dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(lsmc_solve_kernel, dim3(pl.live_of_pass[j]), dim3(kBlock), 0, s, a);
  }
  FSMC_HIP(hipGetLastError());
  double* out = (double*)(ws + pl.off_out);
  hipLaunchKernelGGL(lsmc_final_kernel<Real>, dim3(pl.n), dim3(kBlock), 0, s, a, out, out + 2 * pl.n);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

}  // namespace fsmc

using namespace fsmc;

struct fsmc_lsmc_plan {
  fsmc_ctx* ctx;
  LsmcPlanImpl impl;
  DeviceBuffer ws;
};

extern "C" {

int fsmc_lsmc_plan_create_sweep(fsmc_ctx* ctx, int n_options, const fsmc_mode* m, const double* S0,
                                const double* K, const double* sigma, const double* r, const double* T,
                                const int* steps, const int* paths, int sweep, fsmc_lsmc_plan** out) {
  FSMC_REQUIRE(ctx && out, "NULL argument");
  *out = nullptr;
  FSMC_HIP(hipSetDevice(ctx->device));
  const int grid = sweep == FSMC_LSMC_SWEEP_STREAMED
                       ? 0
                       : resident_grid(ctx->device, m && m->precision == FSMC_FP64 ? FSMC_FP64 : FSMC_FP32);
  fsmc_lsmc_plan* pl = new fsmc_lsmc_plan();
  pl->ctx = ctx;
  int st = lsmc_make_plan(n_options, m, S0, K, sigma, r, T, steps, paths, sweep, grid, &pl->impl);
  if (st == FSMC_OK) {
    std::lock_guard<std::mutex> lock(ctx->mu);
    st = hipSetDevice(ctx->device) == hipSuccess ? FSMC_OK : FSMC_ERR_DEVICE;
    if (st == FSMC_OK) st = pl->ws.reserve(pl->impl.ws);
    if (st == FSMC_OK) {
      char* ws = (char*)pl->ws.ptr;
      const LsmcPlanImpl& q = pl->impl;
      const bool res = q.sweep == FSMC_LSMC_SWEEP_RESIDENT;
      if (hipMemcpyAsync(ws + q.off_meta, q.meta.data(), sizeof(OptMeta) * n_options, hipMemcpyHostToDevice,
                         ctx->stream) != hipSuccess ||
          (!res && hipMemcpyAsync(ws + q.off_table, q.table.data(), sizeof(int4) * q.table.size(),
                                  hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
          (!res && hipMemcpyAsync(ws + q.off_order, q.order.data(), sizeof(int) * n_options, hipMemcpyHostToDevice,
                                  ctx->stream) != hipSuccess) ||
          (res && hipMemcpyAsync(ws + q.off_slots, q.slot_begin.data(), sizeof(int) * q.slot_begin.size(),
                                 hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
          (res && !q.tasks.empty() &&
           hipMemcpyAsync(ws + q.off_tasks, q.tasks.data(), sizeof(int2) * q.tasks.size(), hipMemcpyHostToDevice,
                          ctx->stream) != hipSuccess) ||
          (res && hipMemcpyAsync(ws + q.off_ropt, q.ropt.data(), sizeof(ResOpt) * q.ropt.size(),
                                 hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
          (res && hipMemcpyAsync(ws + q.off_dates, q.dates.data(), sizeof(DateConst) * q.dates.size(),
                                 hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
          hipStreamSynchronize(ctx->stream) != hipSuccess) {
        set_error("upload of the LSMC plan failed");
        st = FSMC_ERR_DEVICE;
      }
    }
  }
  if (st != FSMC_OK) {
    pl->ws.release();
    delete pl;
    return st;
  }
  *out = pl;
  return FSMC_OK;
}

int fsmc_lsmc_plan_create(fsmc_ctx* ctx, int n_options, const fsmc_mode* m, const double* S0, const double* K,
                          const double* sigma, const double* r, const double* T, const int* steps,
                          const int* paths, fsmc_lsmc_plan** out) {
  return fsmc_lsmc_plan_create_sweep(ctx, n_options, m, S0, K, sigma, r, T, steps, paths, FSMC_LSMC_SWEEP_AUTO, out);
}

size_t fsmc_lsmc_plan_workspace(const fsmc_lsmc_plan* pl) { return pl ? pl->impl.ws : 0; }

int fsmc_lsmc_plan_sweep(const fsmc_lsmc_plan* pl) { return pl ? pl->impl.sweep : 0; }

int fsmc_lsmc_plan_run(fsmc_lsmc_plan* pl, void* stream) {
  FSMC_REQUIRE(pl != nullptr, "plan is NULL");
  FSMC_HIP(hipSetDevice(pl->ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : pl->ctx->stream;
  char* ws = (char*)pl->ws.ptr;
  return pl->impl.precision == FSMC_FP64 ? lsmc_enqueue<double>(pl->impl, ws, s)
                                         : lsmc_enqueue<float>(pl->impl, ws, s);
}

int fsmc_lsmc_plan_results(fsmc_lsmc_plan* pl, void* stream, double* price, double* european) {
  FSMC_REQUIRE(pl && price, "NULL argument");
  FSMC_HIP(hipSetDevice(pl->ctx->device));
  hipStream_t s = stream ? (hipStream_t)stream : pl->ctx->stream;
  const double* out = (const double*)((char*)pl->ws.ptr + pl->impl.off_out);
  const int n = pl->impl.n;
  FSMC_HIP(hipMemcpyAsync(price, out, 2 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (european) FSMC_HIP(hipMemcpyAsync(european, out + 2 * n, n * sizeof(double), hipMemcpyDeviceToHost, s));
  int aborted = 0;
  if (pl->impl.sweep == FSMC_LSMC_SWEEP_RESIDENT)
    FSMC_HIP(hipMemcpyAsync(&aborted, (char*)pl->ws.ptr + pl->impl.off_abort, sizeof(int), hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  if (aborted) {
    set_error("resident LSMC sweep: a workgroup gave up waiting at a date barrier");
    return FSMC_ERR_DEVICE;
  }
  return FSMC_OK;
}

int fsmc_lsmc_plan_destroy(fsmc_lsmc_plan* pl) {
  if (!pl) return FSMC_OK;
  (void)hipSetDevice(pl->ctx->device);
  (void)hipDeviceSynchronize();
  pl->ws.release();
  delete pl;
  return FSMC_OK;
}

int fsmc_american_lsmc(fsmc_ctx* ctx, int n_options, const fsmc_mode* m, const double* S0, const double* K,
                       const double* sigma, const double* r, const double* T, const int* steps, const int* paths,
                       double* price, double* european) {
  fsmc_lsmc_plan* pl = nullptr;
  FSMC_TRY(fsmc_lsmc_plan_create(ctx, n_options, m, S0, K, sigma, r, T, steps, paths, &pl));
  int st = fsmc_lsmc_plan_run(pl, nullptr);
  if (st == FSMC_OK) st = fsmc_lsmc_plan_results(pl, nullptr, price, european);
  fsmc_lsmc_plan_destroy(pl);
  return st;
}

int fsmc_american_lsmc_ref(fsmc_ctx* ctx, const fsmc_lsmc_ref_params* p, const fsmc_mode* m, const double* z,
                           double* american, double* european) {
  FSMC_REQUIRE(ctx && p && american && european, "NULL argument");
  FSMC_TRY(validate_mode(m, true, true));
  FSMC_REQUIRE(p->n_paths >= 2 && p->n_steps >= 2 && p->n_paths <= (1 << 30),
               "n_paths %d / n_steps %d out of range", p->n_paths, p->n_steps);
  const bool mrg = m->rng == FSMC_RNG_MRG32K3A;
  const bool buffer = m->rng == FSMC_RNG_BUFFER || mrg;  // both read z from the normals slot
  FSMC_REQUIRE(m->rng != FSMC_RNG_BUFFER || z, "buffer mode needs z");
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t N = (size_t)p->n_paths;
  const size_t real = m->precision == FSMC_FP64 ? 8 : 4;
  const int nb = (int)std::min<size_t>((N + kBlock - 1) / kBlock, 1024);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = align_up(off + bytes, 256);
    return at;
  };
  const size_t o_a = take(real * (size_t)(p->n_steps + 1) * N), o_cf = take(real * N),
               o_part = take(sizeof(double) * 8 * nb), o_beta = take(sizeof(double) * 6), o_out = take(sizeof(double) * 2);
  // MRG32k3a: stream s-1 is the seed stream skipped ahead by nb (s-1)
  // (main.cpp:93-106) and draws n_paths per pricing (utilities.cpp:52-72), so
  // step s of path p reads position nb (s-1) + p of one contiguous run
  const long long nb_skip = (long long)(N / 1024);
  const size_t zn = mrg ? N + (size_t)nb_skip * (size_t)(p->n_steps - 1) : buffer ? N * (size_t)p->n_steps : 0;
  FSMC_TRY(ctx_prepare(ctx, 0, zn * sizeof(double), 0, off));
  hipStream_t s = ctx->stream;
  if (mrg)
    FSMC_TRY(mrg_gaussian_enqueue((uint32_t)m->seed, (uint64_t)m->stream * N, (long long)zn, 0.0, 1.0,
                                  (double*)ctx->normals.ptr, s));
  else if (buffer)
    FSMC_HIP(hipMemcpyAsync(ctx->normals.ptr, z, zn * sizeof(double), hipMemcpyHostToDevice, s));
  char* ws = (char*)ctx->workspace.ptr;
  RefArgs r;
  r.n_paths = p->n_paths;
  r.n_steps = p->n_steps;
  r.spot = p->spot;
  r.strike = p->strike;
  // main.cpp:55-59
  const double dt = p->maturity / p->n_steps;
  r.vsdt_l2e = p->vol * sqrt(dt) * kLog2e;
  r.fwd_l2e = p->vol * p->vol * dt * -0.5 * kLog2e;
  r.z = (const double*)ctx->normals.ptr;
  r.zrow = mrg ? nb_skip : (long long)N;
  r.k0 = (uint32_t)m->seed;
  r.k1 = (uint32_t)(m->seed >> 32);
  r.w = counter_w(kTagLsmc, m->stream);
  r.a = ws + o_a;
  r.cf = ws + o_cf;
  r.partials = (double*)(ws + o_part);
  r.n_partials = nb;
  r.beta = (double*)(ws + o_beta);
  r.out = (double*)(ws + o_out);
  hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(64), 0, s, r.beta, 6);
  const int pb = (int)((N + kBlock - 1) / kBlock);
  const bool f64 = m->precision == FSMC_FP64;
  if (f64 && buffer)
    hipLaunchKernelGGL((ref_paths_kernel<double, true>), dim3(pb), dim3(kBlock), 0, s, r);
  else if (f64)
    hipLaunchKernelGGL((ref_paths_kernel<double, false>), dim3(pb), dim3(kBlock), 0, s, r);
  else if (buffer)
    hipLaunchKernelGGL((ref_paths_kernel<float, true>), dim3(pb), dim3(kBlock), 0, s, r);
  else
    hipLaunchKernelGGL((ref_paths_kernel<float, false>), dim3(pb), dim3(kBlock), 0, s, r);
  FSMC_HIP(hipGetLastError());
  for (int st = p->n_steps - 1; st > 0; --st) {
    if (f64)
      hipLaunchKernelGGL(ref_step_kernel<double>, dim3(nb), dim3(kBlock), 0, s, r, st);
    else
      hipLaunchKernelGGL(ref_step_kernel<float>, dim3(nb), dim3(kBlock), 0, s, r, st);
    hipLaunchKernelGGL(ref_solve_kernel, dim3(1), dim3(kBlock), 0, s, r, st);
  }
  FSMC_HIP(hipGetLastError());
  if (f64)
    hipLaunchKernelGGL(ref_final_kernel<double>, dim3(1), dim3(kBlock), 0, s, r);
  else
    hipLaunchKernelGGL(ref_final_kernel<float>, dim3(1), dim3(kBlock), 0, s, r);
  FSMC_HIP(hipGetLastError());
  double res[2];
  FSMC_HIP(hipMemcpyAsync(res, r.out, sizeof(res), hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  *american = res[0];
  *european = res[1];
  return FSMC_OK;
}

int fsmc_american_paths_ref(fsmc_ctx* ctx, int n_paths, int n_steps, double spot, double fwd_log2, double vsdt_log2,
                            const fsmc_mode* m, const double* z, void* const* rows) {
  FSMC_REQUIRE(ctx && rows, "NULL argument");
  for (int i = 1; i <= n_steps; ++i) FSMC_REQUIRE(rows[i], "rows[%d] is NULL", i);
  FSMC_TRY(validate_mode(m, true, true));
  FSMC_REQUIRE(n_paths >= 1 && n_steps >= 1 && n_paths <= (1 << 30), "n_paths %d / n_steps %d out of range",
               n_paths, n_steps);
  const bool mrg = m->rng == FSMC_RNG_MRG32K3A;
  const bool buffer = m->rng == FSMC_RNG_BUFFER || mrg;
  FSMC_REQUIRE(m->rng != FSMC_RNG_BUFFER || z, "buffer mode needs z");
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t N = (size_t)n_paths;
  const bool f64 = m->precision == FSMC_FP64;
  const size_t elems = (size_t)(n_steps + 1) * N;
  // the same stream layout as fsmc_american_lsmc_ref
  const long long nb_skip = (long long)(N / 1024);
  const size_t zn = mrg ? N + (size_t)nb_skip * (size_t)(n_steps - 1) : buffer ? N * (size_t)n_steps : 0;
  FSMC_TRY(ctx_prepare(ctx, 0, zn * sizeof(double), 0, elems * (f64 ? 8 : 4)));
  hipStream_t s = ctx->stream;
  if (mrg)
    FSMC_TRY(mrg_gaussian_enqueue((uint32_t)m->seed, (uint64_t)m->stream * N, (long long)zn, 0.0, 1.0,
                                  (double*)ctx->normals.ptr, s));
  else if (buffer)
    FSMC_HIP(hipMemcpyAsync(ctx->normals.ptr, z, zn * sizeof(double), hipMemcpyHostToDevice, s));
  RefArgs r{};
  r.n_paths = n_paths;
  r.n_steps = n_steps;
  r.spot = spot;
  r.fwd_l2e = fwd_log2;
  r.vsdt_l2e = vsdt_log2;
  r.z = (const double*)ctx->normals.ptr;
  r.zrow = mrg ? nb_skip : (long long)N;
  r.k0 = (uint32_t)m->seed;
  r.k1 = (uint32_t)(m->seed >> 32);
  r.w = counter_w(kTagLsmc, m->stream);
  r.a = ctx->workspace.ptr;
  const int pb = (int)((N + kBlock - 1) / kBlock);
  if (f64 && buffer)
    hipLaunchKernelGGL((ref_paths_kernel<double, true>), dim3(pb), dim3(kBlock), 0, s, r);
  else if (f64)
    hipLaunchKernelGGL((ref_paths_kernel<double, false>), dim3(pb), dim3(kBlock), 0, s, r);
  else if (buffer)
    hipLaunchKernelGGL((ref_paths_kernel<float, true>), dim3(pb), dim3(kBlock), 0, s, r);
  else
    hipLaunchKernelGGL((ref_paths_kernel<float, false>), dim3(pb), dim3(kBlock), 0, s, r);
  FSMC_HIP(hipGetLastError());
  // row 0 (the spot) stays the caller's, as perform_sims leaves a[0]
  const size_t row_bytes = N * (f64 ? 8 : 4);
  for (int i = 1; i <= n_steps; ++i)
    FSMC_HIP(hipMemcpyAsync(rows[i], (const char*)ctx->workspace.ptr + (size_t)i * row_bytes, row_bytes,
                            hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

}  // extern "C"

// MKL-compatible MT19937 stream with Box-Muller normals on the GPU: the
// generator the AsianOptionsPricing sample draws from
// (AsianOptionsPricing/mkl_rng.h:5-69, MCAsianOptions.cpp:92-93,134-135:
// vslNewStream(VSL_BRNG_MT19937, seed) and vdRngGaussian with
// VSL_RNG_METHOD_GAUSSIAN_BOXMULLER2 on consecutive 8192-normal blocks).
//
// MT19937 (Matsumoto and Nishimura, ACM TOMACS 8(1), 1998): 624-word state,
// seeded here by init_by_array({seed}) (that is what MKL's single-seed
// MT19937 does: its first outputs for seed 42 are 2746317213, 478163327,
// 107420369), one 32-bit tempered output per state word, the whole state
// regenerated ("twisted") every 624 outputs.  Uniforms are u = bits * 2^-32;
// Box-Muller pairs (u1, u2) of consecutive uniforms give
//   z1 = sqrt(-2 ln u1) sin(2 pi u2),  z2 = sqrt(-2 ln u1) cos(2 pi u2)
// (u1 = 0 is taken as 2^-33).  The raw outputs equal MKL's bit for bit; MKL's
// own Box-Muller arithmetic is not reproducible exactly (its normals differ
// from this formula by up to ~2e-8), so the normals match MKL's to that.
//
// The stream is sequential.  A generator object keeps the state in device
// memory; each call continues it, as successive vdRngGaussian calls on one
// VSL stream do.  The bits kernel is one workgroup: the state sits in LDS,
// double-buffered, and the twist runs as three dependent phases (words
// [0, 227) read old words i + 397; [227, 454) read the words phase one
// wrote; [454, 624) those of phase two, the last word also the new word 0),
// each reading the old buffer, writing the new one and ending in one
// barrier.  The Box-Muller transform then runs over the bits in a grid-wide
// kernel.
#include <algorithm>
#include <map>
#include <vector>

#include "../host/gf2_jump.hpp"
#include "fsmc_device.hpp"
#include "fsmc_host.hpp"

namespace fsmc {
namespace {

constexpr int kN = 624, kM = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu, kUpper = 0x80000000u, kLower = 0x7fffffffu;
constexpr int kBitsThreads = 256;

struct MtState {
  uint32_t mt[kN];
  uint32_t pos;  // next word to temper; kN = the state must be twisted first
};

__host__ __device__ __forceinline__ uint32_t twist_word(uint32_t a, uint32_t b, uint32_t far) {
  const uint32_t y = (a & kUpper) | (b & kLower);
  return far ^ (y >> 1) ^ ((y & 1u) ? kMatrixA : 0u);
}

__host__ __device__ __forceinline__ uint32_t temper(uint32_t y) {
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}

// init_by_array({seed}) of the reference implementation.
void init_by_seed(uint32_t seed, MtState& s) {
  uint32_t* mt = s.mt;
  mt[0] = 19650218u;
  for (int i = 1; i < kN; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  int i = 1;
  for (int k = kN; k > 0; --k) {  // key length 1: j stays 0
    mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + seed;
    if (++i >= kN) {
      mt[0] = mt[kN - 1];
      i = 1;
    }
  }
  for (int k = kN - 1; k > 0; --k) {
    mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
    if (++i >= kN) {
      mt[0] = mt[kN - 1];
      i = 1;
    }
  }
  mt[0] = 0x80000000u;
  s.pos = kN;
}

// Next n outputs of the stream in st into out[0, n); one workgroup.  The
// state is double-buffered in LDS, so each twist phase reads the old words
// of one buffer and writes the new words of the other with a single barrier
// per phase; a block that is output whole is tempered and stored straight
// from the registers of the phase that made it.
__global__ __launch_bounds__(kBitsThreads) void mt_bits_kernel(MtState* st, long long n, uint32_t* out) {
  __shared__ uint32_t buf[2][kN];
  const int t = threadIdx.x;
  uint32_t* A = buf[0];  // current state
  uint32_t* B = buf[1];  // the next twist's state
  for (int i = t; i < kN; i += kBitsThreads) A[i] = st->mt[i];
  int pos = (int)st->pos;  // uniform
  __syncthreads();
  long long done = 0;
  constexpr int kH = kN - kM;  // 227
  while (done < n) {
    if (pos == kN) {
      const bool whole = n - done >= kN;  // uniform: the new block goes out entirely
      uint32_t* o = out + done;
      // phase 1: words [0, 227) from old words i, i + 1, i + 397
      if (t < kH) {
        const uint32_t v = twist_word(A[t], A[t + 1], A[t + kM]);
        B[t] = v;
        if (whole) o[t] = temper(v);
      }
      __syncthreads();
      // phase 2: words [227, 454) from old i, i + 1 and new i - 227
      if (t < kH) {
        const int i = t + kH;
        const uint32_t v = twist_word(A[i], A[i + 1], B[i - kH]);
        B[i] = v;
        if (whole) o[i] = temper(v);
      }
      __syncthreads();
      // phase 3: words [454, 624) from old i, new i - 227, and for the last
      // word the new word 0
      if (t + 2 * kH < kN) {
        const int i = t + 2 * kH;
        const uint32_t v = twist_word(A[i], i + 1 < kN ? A[i + 1] : B[0], B[i - kH]);
        B[i] = v;
        if (whole) o[i] = temper(v);
      }
      __syncthreads();
      uint32_t* tmp = A;
      A = B;
      B = tmp;
      pos = 0;
      if (whole) {
        pos = kN;
        done += kN;
        continue;
      }
    }
    const int m = (int)std::min<long long>(kN - pos, n - done);
    for (int i = t; i < m; i += kBitsThreads) out[done + i] = temper(A[pos + i]);
    pos += m;
    done += m;
  }
  __syncthreads();
  for (int i = t; i < kN; i += kBitsThreads) st->mt[i] = A[i];
  if (t == 0) st->pos = (uint32_t)pos;
}

// ---- skip-ahead: one stream over many workgroups ----------------------------
// The raw words X_s of the stream obey X_{s+624} = twist(X_s, X_{s+1},
// X_{s+397}); output s is temper(X_{s+624}).  Let U be the window of the 624
// raw words of the next 624 outputs (U[i] belongs to output P + i, P the
// stream position).  The window J words later is, word by word,
// W[t] = XOR over the set coefficients i of x^J mod phi of X'_{i + t}, with
// X' the sequence continued from U (tests/test_jump_poly.py checks the
// identity on the oracle's stream).  Substream k of a call starts
// J = k L words after U, so workgroup k regenerates X' into an LDS ring and
// correlates it with the coefficient list of x^(kL) mod phi (host-computed,
// sorted), piece by piece: piece p holds the coefficients in
// [p kPiece, (p + 1) kPiece), short enough that X'_{i .. i + 623} of all of
// them is in the ring at once.  Each piece's coefficients are first copied
// to LDS (coalesced), then every thread reads them eight at a time (one
// broadcast 16-byte read) and XORs the ring words they select, so the loop
// has no global-memory wait per coefficient.  The workgroup then emits its
// L outputs from W.
//
// The ring advances 623 words per barrier: X'_s depends on X'_{s-624},
// X'_{s-623} and X'_{s-227}, so thread t < 227 computes s = G + t, then
// s + 227 (from its own word s) and, for t < 169, s + 454 -- every other
// operand is older than G.  (Word G + 623 would need X'_G, another thread's
// word of the same phase.)
constexpr int kJumpThreads = 640;  // thread t < 624 owns window word t
constexpr int kRing = 4096;        // LDS ring of raw words (16 KB)
constexpr int kRingMask = kRing - 1;
constexpr int kLag = kN - kM;      // 227 threads generate
constexpr int kPhase = kN - 1;     // 623 words per phase
constexpr int kPiece = kRing - kN - kPhase;            // 2849 coefficient values per piece
constexpr int kPieces = (19937 + kPiece - 1) / kPiece;  // 7 pieces cover x^0 .. x^19936

// X[s & mask] for s in [g, g + 623), then a barrier; with out, output s
// (temper of X'_s) is also stored for s < m.  Every ring operand of the
// phase is read before any word of it is written (they are all older than
// g), so a phase waits on one round of LDS reads.
__device__ __forceinline__ void mt_ring_phase(uint32_t* X, long long g, uint32_t* out, long long m) {
  const int t = threadIdx.x;
  if (t < kLag) {
    const long long s = g + t;
    const bool third = t < kPhase - 2 * kLag;  // t < 169
    const long long s3 = third ? s + 2 * kLag : s;  // a valid index either way
    const uint32_t a1 = X[(s - kN) & kRingMask], b1 = X[(s - kN + 1) & kRingMask], c1 = X[(s - kLag) & kRingMask];
    const uint32_t a2 = X[(s + kLag - kN) & kRingMask], b2 = X[(s + kLag - kN + 1) & kRingMask];
    const uint32_t a3 = X[(s3 - kN) & kRingMask], b3 = X[(s3 - kN + 1) & kRingMask];
    const uint32_t v1 = twist_word(a1, b1, c1);
    const uint32_t v2 = twist_word(a2, b2, v1);
    const uint32_t v3 = twist_word(a3, b3, v2);
    X[s & kRingMask] = v1;
    X[(s + kLag) & kRingMask] = v2;
    if (third) X[s3 & kRingMask] = v3;
    if (out) {
      if (s < m) out[s] = temper(v1);
      if (s + kLag < m) out[s + kLag] = temper(v2);
      if (third && s3 < m) out[s3] = temper(v3);
    }
  }
  __syncthreads();
}

// XOR of X[(c + t) & mask] over the n coefficients c in C (16-byte aligned).
__device__ __forceinline__ uint32_t ring_correlate(const uint32_t* X, const uint16_t* C, int n, int t) {
  uint32_t a0 = 0, a1 = 0;
  const uint4* C8 = reinterpret_cast<const uint4*>(C);
  int j = 0;
  for (; j + 8 <= n; j += 8) {
    const uint4 v = C8[j >> 3];
    a0 ^= X[((v.x & 0xffffu) + t) & kRingMask] ^ X[((v.x >> 16) + t) & kRingMask];
    a1 ^= X[((v.y & 0xffffu) + t) & kRingMask] ^ X[((v.y >> 16) + t) & kRingMask];
    a0 ^= X[((v.z & 0xffffu) + t) & kRingMask] ^ X[((v.z >> 16) + t) & kRingMask];
    a1 ^= X[((v.w & 0xffffu) + t) & kRingMask] ^ X[((v.w >> 16) + t) & kRingMask];
  }
  for (; j < n; ++j) a0 ^= X[(C[j] + t) & kRingMask];
  return a0 ^ a1;
}

// The window starting at the next output's raw word, from the state (mt, pos):
// mt[pos..623] followed by the first pos words of the next block.
__global__ __launch_bounds__(kBitsThreads) void mt_align_kernel(const MtState* st, uint32_t* U) {
  __shared__ uint32_t A[kN], B[kN];
  const int t = threadIdx.x;
  for (int i = t; i < kN; i += kBitsThreads) A[i] = st->mt[i];
  const int pos = (int)st->pos;
  __syncthreads();
  // the next block, in three dependent phases (as mt_bits_kernel)
  if (t < kLag) B[t] = twist_word(A[t], A[t + 1], A[t + kM]);
  __syncthreads();
  if (t < kLag) B[t + kLag] = twist_word(A[t + kLag], A[t + kLag + 1], B[t]);
  __syncthreads();
  if (t + 2 * kLag < kN) {
    const int i = t + 2 * kLag;
    B[i] = twist_word(A[i], i + 1 < kN ? A[i + 1] : B[0], B[i - kLag]);
  }
  __syncthreads();
  for (int i = t; i < kN; i += kBitsThreads) U[i] = pos + i < kN ? A[pos + i] : B[pos + i - kN];
}

// Workgroup k emits outputs [k L, min(n, (k + 1) L)) of the call into out;
// the last one also leaves the stream's state after output n - 1 in *st.
// The coefficients of x^(k L) mod phi (values < 19937, ascending) are
// poly[poff[(k - 1) (kPieces + 1)] ..): piece p of them is
// poly[poff[(k - 1) (kPieces + 1) + p] .. poff[(k - 1) (kPieces + 1) + p + 1]).
__global__ __launch_bounds__(kJumpThreads) void mt_jump_bits_kernel(const uint32_t* U, const uint16_t* poly,
                                                                     const int* poff, long long L, long long n,
                                                                     uint32_t* out, MtState* st) {
  __shared__ uint32_t X[kRing];
  __shared__ __align__(16) uint16_t C[kPiece + 8];
  const int t = threadIdx.x;
  const int k = blockIdx.x;
  for (int i = t; i < kN; i += kJumpThreads) X[i] = U[i];
  __syncthreads();
  if (k > 0) {
    // W[t] = XOR of X'_{i + t} over the coefficients i
    uint32_t acc = 0;
    long long G = kN;  // X'_s is in the ring for s < G
    const int* po = poff + (size_t)(k - 1) * (kPieces + 1);
    for (int p = 0; p < kPieces; ++p) {
      const int q0 = po[p], nq = po[p + 1] - q0;  // uniform
      if (nq == 0) continue;
      for (int i = t; i < nq; i += kJumpThreads) C[i] = poly[q0 + i];
      // X'_s for s < (p + 1) kPiece + 623 is needed; the oldest word still read is p kPiece
      while (G < (long long)(p + 1) * kPiece + kN) {
        mt_ring_phase(X, G, nullptr, 0);
        G += kPhase;
      }
      __syncthreads();  // C staged (and the last phase written)
      if (t < kN) acc ^= ring_correlate(X, C, nq, t);
      __syncthreads();  // this piece's reads of X and C before either moves on
    }
    if (t < kN) X[t] = acc;
    __syncthreads();
  }
  const long long first = (long long)k * L;
  const long long m = first + L < n ? L : n - first;  // outputs of this substream
  uint32_t* o = out + first;
  for (int i = t; i < kN && i < m; i += kJumpThreads) o[i] = temper(X[i]);
  long long G = kN;
  while (G < m) {  // uniform
    mt_ring_phase(X, G, o, m);
    G += kPhase;
  }
  if (k == (int)gridDim.x - 1) {
    // the state after output n - 1: mt = X'_{m - 624 .. m - 1} with pos = 624
    // (twist first), or the start window with pos = m for a short substream
    __syncthreads();
    const long long s0 = m >= kN ? m - kN : 0;
    for (int i = t; i < kN; i += kJumpThreads) st->mt[i] = X[(s0 + i) & kRingMask];
    if (t == 0) st->pos = (uint32_t)(m >= kN ? kN : m);
  }
}

// Box-Muller pairs over consecutive outputs (even count).
__global__ __launch_bounds__(256) void mt_boxmuller_kernel(const uint32_t* bits, long long pairs, double mean,
                                                           double stddev, double* out) {
  for (long long j = (long long)blockIdx.x * 256 + threadIdx.x; j < pairs; j += (long long)gridDim.x * 256) {
    const uint32_t a = bits[2 * j], b = bits[2 * j + 1];
    const double u1 = a ? (double)a * 0x1.0p-32 : 0x1.0p-33;
    const double u2 = (double)b * 0x1.0p-32;
    const double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    out[2 * j] = mean + stddev * (r * s);
    out[2 * j + 1] = mean + stddev * (r * c);
  }
}

}  // namespace
}  // namespace fsmc

using namespace fsmc;

struct fsmc_mt19937 {
  fsmc_ctx* ctx;
  MtState* state;         // device
  DeviceBuffer bits;      // scratch of the normals calls
  DeviceBuffer jump;      // skip-ahead: start window U, coefficient lists, offsets
  long long jump_L = -1;  // substream length and count the lists in `jump` are for
  int jump_K = 0;
};

namespace {

// A call of at least kJumpMin outputs is split into K substreams of at least
// kPerSub outputs (K <= kMaxSub), each started by a skip-ahead.
constexpr long long kPerSub = 624LL * 64;
constexpr long long kJumpMin = 2 * kPerSub;
constexpr int kMaxSub = 256;

// Sorted coefficient lists of x^(k L) mod phi, k = 1 .. K - 1, with the
// offsets of their pieces, cached per (L, K) for the process (the streams of
// every seed share them).
struct JumpLists {
  std::vector<uint16_t> coef;
  std::vector<int> poff;  // (K - 1) (kPieces + 1) entries: piece p of list k - 1 starts at poff[(k - 1) (kPieces + 1) + p]
};
std::mutex g_jump_mu;
std::map<std::pair<long long, int>, JumpLists>& jump_cache() {
  static std::map<std::pair<long long, int>, JumpLists> c;
  return c;
}

int jump_lists(long long L, int K, const JumpLists** out) {
  std::lock_guard<std::mutex> lock(g_jump_mu);
  auto key = std::make_pair(L, K);
  auto it = jump_cache().find(key);
  if (it == jump_cache().end()) {
    JumpLists jl;
    const std::vector<uint64_t> step = gf2::xpow(gf2::kMT19937, (unsigned long long)L);
    FSMC_REQUIRE(step.size() == (size_t)gf2::kWords, "MT19937 jump polynomial unavailable");
    std::vector<uint64_t> p = step;
    for (int k = 1; k < K; ++k) {
      if (k > 1) p = gf2::mulmod(gf2::kMT19937, p, step);
      int piece = 0;
      for (int w = 0; w < gf2::kWords; ++w)
        for (uint64_t v = p[w]; v; v &= v - 1) {
          const int c = 64 * w + __builtin_ctzll(v);
          while (piece <= c / kPiece) {
            jl.poff.push_back((int)jl.coef.size());
            ++piece;
          }
          jl.coef.push_back((uint16_t)c);
        }
      while (piece <= kPieces) {
        jl.poff.push_back((int)jl.coef.size());
        ++piece;
      }
    }
    it = jump_cache().emplace(key, std::move(jl)).first;
  }
  *out = &it->second;
  return FSMC_OK;
}

// n outputs by K substreams: the start window, then one workgroup per
// substream (skip-ahead, then its outputs); the last leaves the state.
int mt_jump_enqueue(fsmc_mt19937* g, long long n, uint32_t* d_out, hipStream_t s) {
  const int K = (int)std::min<long long>(kMaxSub, n / kPerSub);
  const long long L = (n + K - 1) / K;
  const JumpLists* jl = nullptr;
  FSMC_TRY(jump_lists(L, K, &jl));
  const size_t u_bytes = align_up(kN * sizeof(uint32_t), 256);
  const size_t off_bytes = align_up(jl->poff.size() * sizeof(int), 256);
  const size_t coef_bytes = jl->coef.size() * sizeof(uint16_t);
  char* base = nullptr;
  if (g->jump_L != L || g->jump_K != K || !g->jump.ptr) {
    FSMC_TRY(g->jump.reserve(u_bytes + off_bytes + coef_bytes + 256));
    base = (char*)g->jump.ptr;
    // the lists are the same for every call of this (L, K): uploaded once, in order on s
    FSMC_HIP(hipMemcpyAsync(base + u_bytes, jl->poff.data(), jl->poff.size() * sizeof(int), hipMemcpyHostToDevice,
                            s));
    FSMC_HIP(hipMemcpyAsync(base + u_bytes + off_bytes, jl->coef.data(), coef_bytes, hipMemcpyHostToDevice, s));
    FSMC_HIP(hipStreamSynchronize(s));  // the host vectors are cached, but keep the copy off the pageable path
    g->jump_L = L;
    g->jump_K = K;
  }
  base = (char*)g->jump.ptr;
  uint32_t* U = (uint32_t*)base;
  hipLaunchKernelGGL(mt_align_kernel, dim3(1), dim3(kBitsThreads), 0, s, g->state, U);
  hipLaunchKernelGGL(mt_jump_bits_kernel, dim3(K), dim3(kJumpThreads), 0, s, U,
                     (const uint16_t*)(base + u_bytes + off_bytes), (const int*)(base + u_bytes), L, n, d_out,
                     g->state);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

int mt_validate(long long n, const void* out) {
  FSMC_REQUIRE(n >= 0 && n <= (1LL << 40), "n %lld outside [0, 2^40]", n);
  FSMC_REQUIRE(n == 0 || out != nullptr, "NULL output");
  return FSMC_OK;
}

int mt_bits_enqueue(fsmc_mt19937* g, long long n, uint32_t* d_out, hipStream_t s) {
  if (n == 0) return FSMC_OK;
  if (n >= kJumpMin) return mt_jump_enqueue(g, n, d_out, s);
  hipLaunchKernelGGL(mt_bits_kernel, dim3(1), dim3(kBitsThreads), 0, s, g->state, n, d_out);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

int mt_gaussian_enqueue(fsmc_mt19937* g, long long n, double mean, double stddev, double* d_out, hipStream_t s) {
  FSMC_REQUIRE(n % 2 == 0, "Box-Muller pairs: n %lld must be even", n);
  if (n == 0) return FSMC_OK;
  FSMC_TRY(g->bits.reserve((size_t)n * sizeof(uint32_t)));
  uint32_t* bits = (uint32_t*)g->bits.ptr;
  FSMC_TRY(mt_bits_enqueue(g, n, bits, s));
  const long long pairs = n / 2;
  const unsigned grid = (unsigned)std::min<long long>((pairs + 255) / 256, 8192);
  hipLaunchKernelGGL(mt_boxmuller_kernel, dim3(grid), dim3(256), 0, s, bits, pairs, mean, stddev, d_out);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

}  // namespace

extern "C" {

int fsmc_mt19937_create(fsmc_ctx* ctx, uint32_t seed, fsmc_mt19937** out) {
  FSMC_REQUIRE(ctx && out, "NULL argument");
  *out = nullptr;
  MtState h;
  init_by_seed(seed, h);
  fsmc_mt19937* g = new fsmc_mt19937();
  g->ctx = ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess || hipMalloc((void**)&g->state, sizeof(MtState)) != hipSuccess) {
    (void)hipGetLastError();
    delete g;
    set_error("MT19937 state allocation failed");
    return FSMC_ERR_DEVICE;
  }
  if (hipMemcpy(g->state, &h, sizeof(MtState), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(g->state);
    delete g;
    set_error("MT19937 state upload failed");
    return FSMC_ERR_DEVICE;
  }
  *out = g;
  return FSMC_OK;
}

int fsmc_mt19937_bits_async(fsmc_mt19937* g, long long n, uint32_t* d_out, void* stream) {
  FSMC_REQUIRE(g != nullptr, "generator is NULL");
  FSMC_TRY(mt_validate(n, d_out));
  return mt_bits_enqueue(g, n, d_out, (hipStream_t)stream);
}

int fsmc_mt19937_gaussian_async(fsmc_mt19937* g, long long n, double mean, double stddev, double* d_out,
                                void* stream) {
  FSMC_REQUIRE(g != nullptr, "generator is NULL");
  FSMC_TRY(mt_validate(n, d_out));
  return mt_gaussian_enqueue(g, n, mean, stddev, d_out, (hipStream_t)stream);
}

int fsmc_mt19937_bits(fsmc_mt19937* g, long long n, uint32_t* out) {
  FSMC_REQUIRE(g != nullptr, "generator is NULL");
  FSMC_TRY(mt_validate(n, out));
  fsmc_ctx* ctx = g->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t bytes = (size_t)n * sizeof(uint32_t);
  FSMC_TRY(ctx_prepare(ctx, 0, bytes, 0, 0));
  FSMC_TRY(mt_bits_enqueue(g, n, (uint32_t*)ctx->normals.ptr, ctx->stream));
  if (n) FSMC_HIP(hipMemcpyAsync(out, ctx->normals.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  FSMC_HIP(hipStreamSynchronize(ctx->stream));
  return FSMC_OK;
}

int fsmc_mt19937_gaussian(fsmc_mt19937* g, long long n, double mean, double stddev, double* out) {
  FSMC_REQUIRE(g != nullptr, "generator is NULL");
  FSMC_TRY(mt_validate(n, out));
  fsmc_ctx* ctx = g->ctx;
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t bytes = (size_t)n * sizeof(double);
  FSMC_TRY(ctx_prepare(ctx, 0, bytes, 0, 0));
  FSMC_TRY(mt_gaussian_enqueue(g, n, mean, stddev, (double*)ctx->normals.ptr, ctx->stream));
  if (n) FSMC_HIP(hipMemcpyAsync(out, ctx->normals.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  FSMC_HIP(hipStreamSynchronize(ctx->stream));
  return FSMC_OK;
}

int fsmc_mt19937_destroy(fsmc_mt19937* g) {
  if (!g) return FSMC_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipDeviceSynchronize();
  if (g->state) (void)hipFree(g->state);
  g->bits.release();
  g->jump.release();
  delete g;
  return FSMC_OK;
}

}  // extern "C"

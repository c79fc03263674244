// Jump-ahead for the F2-linear sample generators (MT19937, SFMT19937).
//
// For a generator whose state moves by a linear map T over GF(2), the
// state J steps ahead is g(T) s with g = x^J mod phi, phi the minimal
// polynomial of T on the generator's period (degree 19937 for both).  With
// s the window of the last N generated words, g(T) s is again such a window:
// word k of it is the XOR of the words at positions i + k over the set
// coefficients g_i (Haramoto, Matsumoto, Nishimura, Saito, L'Ecuyer,
// "Efficient jump ahead for F2-linear random number generators", INFORMS J.
// Computing 20(3), 2008).  This header is the host side: the polynomials.
#pragma once

#include <cstdint>
#include <vector>

namespace fsmc {
namespace gf2 {

enum Gen : int { kMT19937 = 0, kSFMT19937 = 1 };

// Degree of phi: 19937 for MT19937 (whose generated words move in the
// 19937-dimensional part of the state), 19968 for SFMT19937 (its 128-bit
// step's characteristic polynomial: a certified state still has a
// component outside the 19937-dimensional period subspace, so the whole
// state is jumped).
constexpr int kMaxDeg = 19968;
constexpr int kWords = kMaxDeg / 64;  // 312 64-bit words hold every reduced polynomial
int degree(Gen g);

// phi of the generator's step (MT19937: one 32-bit word; SFMT19937: one
// 128-bit word), found once per process by Berlekamp-Massey on the low bit
// of 2 x 19968 + 128 successive words of a host copy of the generator;
// degree(g) + 1 coefficients, lowest first; empty if the degree found is not
// degree(g).
const std::vector<uint64_t>& minpoly(Gen g);

// x^J mod phi as kWords words (coefficient i in bit i % 64 of word i / 64);
// empty if phi is not available.
std::vector<uint64_t> xpow(Gen g, unsigned long long J);
// a * b mod phi (both reduced).
std::vector<uint64_t> mulmod(Gen g, const std::vector<uint64_t>& a, const std::vector<uint64_t>& b);

// Host reference generators (used for Berlekamp-Massey and by the tests):
// the raw (untempered) word sequence.  MT19937: words X_624, X_625, ... of
// init_by_array({seed}); SFMT19937: 32-bit lanes of the 128-bit words of
// the state regenerations, in output order.
std::vector<uint32_t> mt19937_raw(uint32_t seed, long long n);
std::vector<uint32_t> sfmt19937_raw(uint32_t seed, long long n);

}  // namespace gf2
}  // namespace fsmc

// mt_jump.hpp -- jump-ahead for std::mt19937 (host code, mt_jump.cpp).
//
// MT19937 is linear over GF(2): its state advances by a fixed linear map
// whose characteristic polynomial phi(t) has degree 19937, so every bit of
// the untempered word sequence x_0, x_1, ... satisfies the recurrence phi
// defines, and for any jump J, with g(t) = t^J mod phi(t) = sum g_i t^i,
//     x_{n+J} = XOR of x_{n+i} over the i with g_i = 1      (every n >= 1)
// (Haramoto, Matsumoto, Nishimura, L'Ecuyer, Panneton, "Efficient jump ahead
// for F2-linear random number generators", INFORMS J. Computing 20(3), 2008).
// The array of 624 words a generator holds before producing block x_{m+624..}
// is (x_m .. x_{m+623}), of which x_m counts by its top bit only, so the
// array for a block J words later is the XOR of the shifted windows.
#pragma once

#include <cstdint>
#include <vector>

namespace fgl {

constexpr int kMtDegree = 19937;
constexpr int kMtPolyWords = 312;   // uint64 words holding the coefficients of degrees < 19968

// out = the polynomials t^(c * step) mod phi(t) for c = 1..n, kMtPolyWords
// words each (bit b of word w: the coefficient of t^(64 w + b)).  phi is
// found once per process (Berlekamp-Massey on the generator's low bits); a
// table per step is kept and extended as larger n are asked for.  Returns
// false when phi does not come out at degree 19937.
bool mt_jump_polys(uint64_t step, int n, std::vector<uint64_t> &out);

}  // namespace fgl

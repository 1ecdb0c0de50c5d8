// rng.h - per-history random streams and the samplers built on them.
//
// Bit-for-bit restatement of what the transport path draws:
//  * pcg32 (O'Neill's PCG-XSH-RR 64/32, output of the pre-advance state),
//    64-bit LCG multiplier 6364136223846793005 and increment
//    1442695040888963407; seeding pcg32(s) sets state = (s + inc) * mult + inc;
//    advance(n) is Brown's (1994) arbitrary-stride jump.  Streams are
//    `pcg32(seed).advance(stride * history_id)`
//    (include/simulation/particle.hpp:188-193, src/simulation.cpp:63-67).
//  * rand(): one double in [0,1) made from TWO consecutive 32-bit outputs x1,
//    x2 as (x1 + x2 * 2^32) / 2^64, clamped to nextafter(1, 0); this is
//    libstdc++'s generate_canonical<double, 53> over a 32-bit engine, which
//    RNG::rand uses (include/utils/rng.hpp:41-43).
//  * exponential(lambda) = -log(1 - rand()) / lambda, DBL_MAX without a draw
//    when lambda == 0 (include/utils/rng.hpp:73-78).
//  * discrete(cdf): no draw when there are fewer than two weights, otherwise
//    the first index i with cdf[i] >= rand() (include/utils/rng.hpp:87-95).
//    The cdf is built once per table on the host by build_discrete_cdf with
//    the same arithmetic the per-call std::discrete_distribution uses.
#pragma once

#include <stdint.h>

#include "detmath.h"

#if defined(__HIPCC__)
#define RNG_HD __host__ __device__ __forceinline__
#else
#define RNG_HD static inline
#endif

#define ABX_PCG_MULT 6364136223846793005ULL
#define ABX_PCG_INC 1442695040888963407ULL
#define ABX_DBL_MAX 1.7976931348623157e308

RNG_HD uint64_t pcg_seed(uint64_t seed) {
  return (seed + ABX_PCG_INC) * ABX_PCG_MULT + ABX_PCG_INC;
}

// Brown's jump-ahead: state after `delta` steps.
RNG_HD uint64_t pcg_advance(uint64_t state, uint64_t delta) {
  uint64_t cur_mult = ABX_PCG_MULT, cur_plus = ABX_PCG_INC;
  uint64_t acc_mult = 1u, acc_plus = 0u;
  while (delta > 0u) {
    if (delta & 1u) {
      acc_mult *= cur_mult;
      acc_plus = acc_plus * cur_mult + cur_plus;
    }
    cur_plus = (cur_mult + 1u) * cur_plus;
    cur_mult *= cur_mult;
    delta >>= 1;
  }
  return acc_mult * state + acc_plus;
}

// Start of history `h`'s stream.
RNG_HD uint64_t pcg_history_state(uint64_t seed, uint64_t stride, uint64_t h) {
  return pcg_advance(pcg_seed(seed), stride * h);
}

RNG_HD uint32_t pcg_next(uint64_t& state) {
  const uint64_t old = state;
  state = old * ABX_PCG_MULT + ABX_PCG_INC;
  const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  const uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

RNG_HD double rng_rand(uint64_t& state) {
  const uint32_t x1 = pcg_next(state);
  const uint32_t x2 = pcg_next(state);
  double sum = 0.0;
  sum += (double)x1 * 1.0;
  sum += (double)x2 * 4294967296.0;
  double r = sum / 18446744073709551616.0;
  if (r >= 1.0) r = 0.99999999999999988898;  // nextafter(1, 0)
  return r;
}

RNG_HD double rng_exponential(uint64_t& state, double lambda) {
  if (lambda == 0.0) return ABX_DBL_MAX;
  const double u = rng_rand(state);
  return -dm_log(1.0 - u) / lambda;
}

// First index i in [0, n) with cdf[i] >= u (std::lower_bound semantics).
RNG_HD int cdf_lower_bound(const double* cdf, int n, double u) {
  int lo = 0, len = n;
  while (len > 0) {
    const int half = len >> 1;
    if (cdf[lo + half] < u) {
      lo = lo + half + 1;
      len = len - half - 1;
    } else {
      len = half;
    }
  }
  return lo;
}

RNG_HD int rng_discrete(uint64_t& state, const double* cdf, int n) {
  if (n < 2) return 0;
  const double u = rng_rand(state);
  return cdf_lower_bound(cdf, n, u);
}

// Host-side: the cumulative table std::discrete_distribution builds from a
// weight vector (accumulate, divide each weight by the sum, partial sums,
// last entry forced to 1).
static inline void build_discrete_cdf(const double* w, int n, double* cdf) {
  if (n < 2) {
    for (int i = 0; i < n; i++) cdf[i] = 1.0;
    return;
  }
  double sum = 0.0;
  for (int i = 0; i < n; i++) sum = sum + w[i];
  double acc = 0.0;
  for (int i = 0; i < n; i++) {
    const double p = w[i] / sum;
    acc = (i == 0) ? p : acc + p;
    cdf[i] = acc;
  }
  cdf[n - 1] = 1.0;
}

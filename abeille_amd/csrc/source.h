// source.h - source-particle sampling for one history (device code).
//
// Simulation::sample_sources (src/simulation.cpp:55-77) and
// Source::generate_particle (src/source.cpp:44-90): stream
// pcg32(seed).advance(stride * h); [source pick if >= 2 sources];
// isotropic direction mu = 2 xi - 1, phi = 2 pi xi (src/isotropic.cpp:28-36);
// mono-energetic energy (no draw); point (no draw) or box position
// (3 draws per attempt, src/box.cpp:36-41), resampled while lost and, for
// fissile-only sources, while not in a fissile material.
#pragma once

#include "../../include/abx.h"
#include "physics.h"

struct DevSources {
  const abx_source* src;
  const double* cdf;  // discrete table over source weights
  int32_t n;
};

ABX_HD V3 source_position(const abx_source& s, uint64_t& rng) {
  if (s.spatial == 0) return V3{s.low[0], s.low[1], s.low[2]};
  const double x = (s.high[0] - s.low[0]) * rng_rand(rng) + s.low[0];
  const double y = (s.high[1] - s.low[1]) * rng_rand(rng) + s.low[1];
  const double z = (s.high[2] - s.low[2]) * rng_rand(rng) + s.low[2];
  return V3{x, y, z};
}

// Returns 0 or an ABX_E_* code.
template <class TK>
ABX_HD int sample_source(const DevProblem& P, const DevSources& S, uint64_t h, TK& t,
                         abx_particle& out) {
  uint64_t rng = pcg_history_state(P.seed, P.stride, h);
  const int idx = S.n >= 2 ? rng_discrete(rng, S.cdf, S.n) : 0;
  const abx_source& s = S.src[idx];
  V3 u;
  if (s.direction == 0) {
    const double mu = 2. * rng_rand(rng) - 1.;
    const double phi = 2. * ABX_PI * rng_rand(rng);
    u = dir_mu_phi(mu, phi);
  } else {
    u = vdir(s.dir[0], s.dir[1], s.dir[2]);
  }
  const double E = s.energy;
  if (E <= P.min_energy || P.max_energy <= E) return -1;  // 200 resamplings fail
  V3 r = source_position(s, rng);
  trk_init(P, t, r, u, 0);
  int guard = 0;
  while (t.cell < 0) {
    if (++guard > 100000 || t.err) return -3;
    r = source_position(s, rng);
    trk_init(P, t, r, u, 0);
  }
  if (s.fissile_only) {
    int count = 0;
    while (t.cell < 0 || !P.fissile[P.cell[t.cell].fill]) {
      if (count == 201 || t.err) return -1;
      r = source_position(s, rng);
      trk_init(P, t, r, u, 0);
      count++;
    }
  }
  out.r[0] = r.x;
  out.r[1] = r.y;
  out.r[2] = r.z;
  out.u[0] = u.x;
  out.u[1] = u.y;
  out.u[2] = u.z;
  out.E = E;
  out.wgt = 1.0;
  out.wgt2 = 0.;
  out.history_id = h;
  out.family_id = h;
  out.rng_state = rng;
  return 0;
}

// physics.h - one history of the multigroup transport loop (device code).
//
// Restates, per particle and with the reference's draw order:
//  * DeltaTracker::transport loop body      src/delta_tracker.cpp:86-229
//  * CarterTracker::transport loop body     src/carter_tracker.cpp:104-262
//  * SurfaceTracker::transport loop body    src/surface_tracker.cpp:70-200
//  * Transporter::collision / branching_collision / make_fission_neutrons /
//    do_scatter / russian_roulette          src/transporter.cpp:35-93,269-487
//  * MGNuclide::sample_scatter / sample_fission, MGAngleDistribution::sample_mu
//                                           src/mg_nuclide.cpp:442-544,
//                                           include/materials/mg_angle_distribution.hpp:45-60
//  * rotate_direction / Direction(mu, phi)  include/utils/direction.hpp:43-62,130-151
//  * Tracker::do_reflection                 include/simulation/tracker.hpp:314-359
//  * CollisionMeshTally::score_collision    src/collision_mesh_tally.cpp:33-110
//  * TrackLengthMeshTally::score_flight     src/track_length_mesh_tally.cpp:79-393
//
// Everything a kernel needs to do with global memory (tally atomics, the
// secondary stack, fission-bank appends, traces) goes through the `Sink`
// template parameter, so the same code runs in the history kernel and in
// the event kernels.
//
// Sink::split selects the split-precision variant: sampled flight
// distances, cross-section ratios, scattering / emission directions and
// tally score factors in FP32 (hardware log / sincos / rcp / sqrt);
// positions, surface and lattice tests, weights, k scores and tally
// accumulators stay FP64.  With split == false every operation is the
// reference's FP64 one, in its order.
#pragma once

#include "detmath.h"
#include "geometry.h"
#include "rng.h"

struct Part {
  V3 r, u;
  V3 rb;  // r_birth
  double E, w, w2;
  uint64_t rng;
  uint64_t hid, fam;
  uint32_t calls;      // 32-bit draws since the history entered transport
  uint32_t daughters;  // Particle::daughter_counter
  int32_t sec_id;
  int32_t g;           // energy group of E
  int32_t alive;
  // Particle::previous_{position,direction,energy}, Esmp_, reflected and
  // previous_collision_virtual_ (include/simulation/particle.hpp:231-239):
  // kept by the extended kernels only (Sink::ext), for the parent records
  // of exact cancellation
  V3 pr, pu;
  double pE, Esmp;
  int32_t pvirt, refl;
};

// Particle::move (particle.hpp:125-133) in the extended kernels: the
// previous position follows every move except the first after a reflection.
template <bool EXT>
ABX_HD void part_move(Part& p, double d) {
  if constexpr (EXT) {
    if (!p.refl)
      p.pr = p.r;
    else
      p.refl = 0;
  }
  p.r = vadd(p.r, vscale(p.u, d));
}

struct SecRec {
  double r[3], u[3], E, w, w2;
};

// RNG with draw counting (for traces / stride checks)
ABX_HD double p_rand(Part& p) {
  p.calls += 2;
  return rng_rand(p.rng);
}
ABX_HD double p_exponential(Part& p, double lambda) {
  if (lambda == 0.0) return ABX_INF;
  p.calls += 2;
  return rng_exponential(p.rng, lambda);
}
ABX_HD int p_discrete(Part& p, const double* cdf, int n) {
  if (n < 2) return 0;
  p.calls += 2;
  return rng_discrete(p.rng, cdf, n);
}

// Flight distance -log(1 - xi) / lambda (RNG::exponential), FP32 in split mode.
template <bool SPLIT>
ABX_HD double flight_distance(Part& p, double lambda) {
  if constexpr (!SPLIT) {
    return p_exponential(p, lambda);
  } else {
    if (lambda == 0.0) return ABX_INF;
    p.calls += 2;
    const double xi = rng_rand(p.rng);
    // For xi below 2^-25, (float)(1 - xi) is 1.0f and the FP32 log would
    // give a flight of zero length (one flight in 3e7): a particle that has
    // just reflected would collide on the reflecting surface and could leave
    // the geometry with its new direction.  Below 2^-12 the series
    // xi + xi^2 / 2 is -log(1 - xi) to FP32 precision.
    const float t = xi < 0x1p-12 ? (float)(xi + 0.5 * xi * xi) : -fp_log((float)(1.0 - xi));
    return (double)fp_div(t, (float)lambda);
  }
}

ABX_HD int energy_index(const DevProblem& P, double E) {
  int i = 0;
  for (i = 0; i < P.G; i++)
    if (P.ebounds[i] <= E && E < P.ebounds[i + 1]) break;
  return i;
}

// Direction(mu, phi) with clamping, then normalisation
ABX_NOINL V3 dir_mu_phi(double mu, double phi) {
  if (mu < -1.)
    mu = -1.;
  else if (mu > 1.)
    mu = 1.;
  if (phi < 0.)
    phi = 0.;
  else if (phi > 2 * ABX_PI)
    phi = 2 * ABX_PI;
  double sn, cs;
  dm_sincos(phi, &sn, &cs);
  const double x = sqrt(1. - mu * mu) * cs;
  const double y = sqrt(1. - mu * mu) * sn;
  return vdir(x, y, mu);
}

ABX_NOINL V3 rotate_dir(const V3& u, double mu, double phi) {
  double sn, cs;
  dm_sincos(phi, &sn, &cs);
  const double sqrt_mu = sqrt(1. - mu * mu);
  const double sqrt_w = sqrt(1. - u.z * u.z);
  double ux, uy, uz;
  if (sqrt_w > 1.E-10) {
    ux = mu * u.x + sqrt_mu * (u.x * u.z * cs - u.y * sn) / sqrt_w;
    uy = mu * u.y + sqrt_mu * (u.y * u.z * cs + u.x * sn) / sqrt_w;
    uz = mu * u.z - sqrt_mu * sqrt_w * cs;
  } else {
    const double sqrt_v = sqrt(1. - u.y * u.y);
    ux = mu * u.x + sqrt_mu * (u.x * u.y * cs + u.z * sn) / sqrt_v;
    uy = mu * u.y - sqrt_mu * sqrt_v * cs;
    uz = mu * u.z + sqrt_mu * (u.y * u.z * cs - u.x * sn) / sqrt_v;
  }
  return vdir(ux, uy, uz);
}

// rotate_direction in FP32, normalised with a reciprocal square root.  The
// branch threshold is the FP32 analogue of the reference's 1e-10 on
// sqrt(1 - w^2); both branches are the same rotation up to the azimuth
// origin, which is uniform.
ABX_HD V3 rotate_dir_fp(const V3& ud, double mu_d, double phi_d) {
  const float mu = (float)mu_d;
  float sn, cs;
  fp_sincos((float)phi_d, &sn, &cs);
  const float ux = (float)ud.x, uy = (float)ud.y, uz = (float)ud.z;
  const float m2 = 1.f - mu * mu;
  const float sqrt_mu = m2 > 0.f ? fp_sqrt(m2) : 0.f;
  const float w2 = 1.f - uz * uz;
  float x, y, z;
  if (w2 > 1e-6f) {
    const float sqrt_w = fp_sqrt(w2);
    const float iw = fp_rcp(sqrt_w);
    x = mu * ux + sqrt_mu * (ux * uz * cs - uy * sn) * iw;
    y = mu * uy + sqrt_mu * (uy * uz * cs + ux * sn) * iw;
    z = mu * uz - sqrt_mu * sqrt_w * cs;
  } else {
    const float sqrt_v = fp_sqrt(1.f - uy * uy);
    const float iv = fp_rcp(sqrt_v);
    x = mu * ux + sqrt_mu * (ux * uy * cs + uz * sn) * iv;
    y = mu * uy - sqrt_mu * sqrt_v * cs;
    z = mu * uz + sqrt_mu * (uy * uz * cs - ux * sn) * iv;
  }
  const float inv = fp_rsqrt(x * x + y * y + z * z);
  return V3{(double)(x * inv), (double)(y * inv), (double)(z * inv)};
}

template <bool SPLIT>
ABX_HD V3 rotate(const V3& u, double mu, double phi) {
  if constexpr (SPLIT)
    return rotate_dir_fp(u, mu, phi);
  else
    return rotate_dir(u, mu, phi);
}

ABX_NOINL double sample_mu(const DevProblem& P, Part& p, int m, int gi, int go) {
  const double xi = p_rand(p);
  const int base = P.ang_base[m];
  if (base < 0) {
    // the isotropic table {mu -1, 1; pdf 0.5, 0.5; cdf 0, 1} evaluated below
    // reduces to this (xi == 0 returns mu[0] = -1, the same value)
    return -1. + ((xi - 0.) / 0.5);
  }
  const int k = base + gi * P.G + go;
  const int b0 = P.ang_off[k], n = P.ang_off[k + 1] - b0;
  const double* mu = P.ang_mu + b0;
  const double* pdf = P.ang_pdf + b0;
  const double* cdf = P.ang_cdf + b0;
  int l = cdf_lower_bound(cdf, n, xi);
  if (xi == cdf[l]) return mu[l];
  l--;
  if (pdf[l] == pdf[l + 1]) return mu[l] + ((xi - cdf[l]) / pdf[l]);
  const double mm = (pdf[l + 1] - pdf[l]) / (mu[l + 1] - mu[l]);
  return mu[l] + (1. / mm) * (sqrt(pdf[l] * pdf[l] + 2. * mm * (xi - cdf[l])) - pdf[l]);
}

ABX_HD int tally_ebin(const DevProblem& P, const DTally& T, int ti, const Part& p) {
  if (p.g < P.G && p.E == P.emid[p.g]) return P.tally_gbin[ti * P.G + p.g];
  const double* eb = P.tally_eb + T.eb_off;
  for (int e = 0; e < T.Ne; e++)
    if (eb[e] <= p.E && p.E <= eb[e + 1]) return e;
  return -1;
}

ABX_HD double quantity_factor(const DevProblem& P, const DTally& T, const Part& p, int mg) {
  switch (T.quantity) {
    case 1: return p.w * P.Eel[mg];
    case 2: return p.w * P.Eabs[mg];
    case 3: return p.w * P.Ef[mg];
    case 4: return p.w * P.Et[mg];
    case 7: return p.w2;
    default: return p.w;
  }
}

template <class Sink>
ABX_NOINL void score_collision_tallies(const DevProblem& P, Sink& s, const Part& p, int m) {
  const int mg = m * P.G + p.g;
  for (int ti = 0; ti < P.ntally; ti++) {
    const DTally& T = P.tally[ti];
    if (T.estimator != 0) continue;
    const double Et = P.Et[mg];
    double scr;
    if constexpr (Sink::split)
      scr = (double)fp_rcp((float)Et * (float)T.net_weight);
    else
      scr = 1. / (Et * T.net_weight);
    const int i = (int)floor((p.r.x - T.lx) * T.dxi);
    const int j = (int)floor((p.r.y - T.ly) * T.dyi);
    const int k = (int)floor((p.r.z - T.lz) * T.dzi);
    const int l = tally_ebin(P, T, ti, p);
    if (l == -1) continue;
    if (i >= 0 && i < T.Nx && j >= 0 && j < T.Ny && k >= 0 && k < T.Nz) {
      scr *= quantity_factor(P, T, p, mg);
      const int64_t idx = (((int64_t)l * T.Nx + i) * T.Ny + j) * T.Nz + k;
      s.tally_add(P.tally_data + T.data_off + idx, scr);
    }
  }
}

#ifndef ABX_TLE_DDA
#define ABX_TLE_DDA 1
#endif

// TrackLengthMeshTally::initialize_indices
ABX_HD void tle_init(const DTally& T, const V3& r, const V3& u, int& i, int& j, int& k,
                     int on[3]) {
  i = (int)floor((r.x - T.lx) * T.dxi);
  j = (int)floor((r.y - T.ly) * T.dyi);
  k = (int)floor((r.z - T.lz) * T.dzi);
  on[0] = on[1] = on[2] = 0;
  const double xc = T.lx + i * T.dx + 0.5 * T.dx;
  const double yc = T.ly + j * T.dy + 0.5 * T.dy;
  const double zc = T.lz + k * T.dz + 0.5 * T.dz;
  const double xl = xc - 0.5 * T.dx, xh = xc + 0.5 * T.dx;
  const double yl = yc - 0.5 * T.dy, yh = yc + 0.5 * T.dy;
  const double zl = zc - 0.5 * T.dz, zh = zc + 0.5 * T.dz;
  if (dabs(xl - r.x) < ABX_SURFACE_COINCIDENT) {
    if (u.x < 0.) { i--; on[0] = 1; } else { on[0] = -1; }
  } else if (dabs(xh - r.x) < ABX_SURFACE_COINCIDENT) {
    if (u.x < 0.) { on[0] = 1; } else { i++; on[0] = -1; }
  }
  if (dabs(yl - r.y) < ABX_SURFACE_COINCIDENT) {
    if (u.y < 0.) { j--; on[1] = 1; } else { on[1] = -1; }
  } else if (dabs(yh - r.y) < ABX_SURFACE_COINCIDENT) {
    if (u.y < 0.) { on[1] = 1; } else { j++; on[1] = -1; }
  }
  if (dabs(zl - r.z) < ABX_SURFACE_COINCIDENT) {
    if (u.z < 0.) { k--; on[2] = 1; } else { on[2] = -1; }
  } else if (dabs(zh - r.z) < ABX_SURFACE_COINCIDENT) {
    if (u.z < 0.) { on[2] = 1; } else { k++; on[2] = -1; }
  }
}

// TrackLengthMeshTally::find_entry_point (slab method)
ABX_HD bool tle_entry(const DTally& T, V3& r, const V3& u, const V3& ui, double& d) {
  const double ux_inv = ui.x, uy_inv = ui.y, uz_inv = ui.z;
  double d_min = (T.lx - r.x) * ux_inv, d_max = (T.hx - r.x) * ux_inv;
  double t;
  if (d_min > d_max) { t = d_min; d_min = d_max; d_max = t; }
  double y0 = (T.ly - r.y) * uy_inv, y1 = (T.hy - r.y) * uy_inv;
  if (y0 > y1) { t = y0; y0 = y1; y1 = t; }
  if ((d_min > y1) || (y0 > d_max)) return false;
  if (y0 > d_min) d_min = y0;
  if (y1 < d_max) d_max = y1;
  double z0 = (T.lz - r.z) * uz_inv, z1 = (T.hz - r.z) * uz_inv;
  if (z0 > z1) { t = z0; z0 = z1; z1 = t; }
  if ((d_min > z1) || (z0 > d_max)) return false;
  if (z0 > d_min) d_min = z0;
  if (z1 < d_max) d_max = z1;
  if (d_max < d_min) { t = d_max; d_max = d_min; d_min = t; }
  if ((d_max < 0.) && (d_min < 0.)) return false;
  if (d_min < 0.) return false;
  r = vadd(r, vscale(u, d_min));
  d -= d_min;
  return true;
}

template <class Sink>
ABX_NOINL void score_flight_tallies(const DevProblem& P, Sink& s, const Part& p, double d,
                                 int m) {
  const int mg = m * P.G + p.g;
  for (int ti = 0; ti < P.ntally; ti++) {
    const DTally& T = P.tally[ti];
    if (T.estimator != 1) continue;
    V3 r = p.r;
    const V3 u = p.u;
    V3 ui;  // 1 / u per axis
    if constexpr (Sink::split)
      ui = V3{(double)fp_rcp((float)u.x), (double)fp_rcp((float)u.y), (double)fp_rcp((float)u.z)};
    else
      ui = V3{1. / u.x, 1. / u.y, 1. / u.z};
    double dd = d;
    int i, j, k, on[3];
    tle_init(T, r, u, i, j, k, on);
    const bool inside = i >= 0 && i < T.Nx && j >= 0 && j < T.Ny && k >= 0 && k < T.Nz;
    if (!inside) {
      if (!tle_entry(T, r, u, ui, dd)) continue;
      tle_init(T, r, u, i, j, k, on);
    }
    double base = 1. / T.net_weight;
    base *= quantity_factor(P, T, p, mg);
    const int l = tally_ebin(P, T, ti, p);
    if (l == -1) continue;
    double remaining = dd;
    double* data = P.tally_data + T.data_off + (int64_t)l * T.Nx * T.Ny * T.Nz;
#if ABX_TLE_DDA
    // The walk of TrackLengthMeshTally::score_flight
    // (src/track_length_mesh_tally.cpp:34-435) as a digital differential
    // analyser: the distance along the flight to the next plane of each
    // axis is advanced by one cell pitch per crossing instead of being
    // recomputed from an advanced position, and the nearest plane wins in
    // the reference's order (x, then y, then z).  Same cells, same order;
    // segment lengths agree to rounding.
    const int sx = u.x > 0. ? 1 : (u.x < 0. ? -1 : 0);
    const int sy = u.y > 0. ? 1 : (u.y < 0. ? -1 : 0);
    const int sz = u.z > 0. ? 1 : (u.z < 0. ? -1 : 0);
    double tx = sx ? (T.lx + (i + (sx > 0)) * T.dx - r.x) * ui.x : ABX_INF;
    double ty = sy ? (T.ly + (j + (sy > 0)) * T.dy - r.y) * ui.y : ABX_INF;
    double tz = sz ? (T.lz + (k + (sz > 0)) * T.dz - r.z) * ui.z : ABX_INF;
    const double px = sx ? T.dx * dabs(ui.x) : 0., py = sy ? T.dy * dabs(ui.y) : 0.,
                 pz = sz ? T.dz * dabs(ui.z) : 0.;
    double t = 0.;
    while (remaining > 0.) {
      int axis = 0;
      double tn = tx;
      if (ty < tn) { tn = ty; axis = 1; }
      if (tz < tn) { tn = tz; axis = 2; }
      if (tn == ABX_INF) break;
      const double step = tn - t;
      const double d_tile = step < remaining ? step : remaining;
      if (i >= 0 && i < T.Nx && j >= 0 && j < T.Ny && k >= 0 && k < T.Nz) {
        s.tally_add(data + ((int64_t)i * T.Ny + j) * T.Nz + k, d_tile * base);
      } else {
        break;
      }
      remaining -= d_tile;
      if (remaining <= 0.) break;
      t = tn;
      if (axis == 0) { i += sx; tx += px; }
      else if (axis == 1) { j += sy; ty += py; }
      else { k += sz; tz += pz; }
    }
  }
}

#else
    while (remaining > 0.) {
      const double xc = T.lx + i * T.dx + 0.5 * T.dx;
      const double yc = T.ly + j * T.dy + 0.5 * T.dy;
      const double zc = T.lz + k * T.dz + 0.5 * T.dz;
      const double rtx = r.x - xc, rty = r.y - yc, rtz = r.z - zc;
      double dist = ABX_INF;
      int key = 0;
      const double diff_xl = -T.dx * 0.5 - rtx, diff_xh = T.dx * 0.5 - rtx;
      const double diff_yl = -T.dy * 0.5 - rty, diff_yh = T.dy * 0.5 - rty;
      const double diff_zl = -T.dz * 0.5 - rtz, diff_zh = T.dz * 0.5 - rtz;
      const double ux_inv = ui.x, uy_inv = ui.y, uz_inv = ui.z;
      const double d_xl = diff_xl * ux_inv, d_xh = diff_xh * ux_inv;
      const double d_yl = diff_yl * uy_inv, d_yh = diff_yh * uy_inv;
      const double d_zl = diff_zl * uz_inv, d_zh = diff_zh * uz_inv;
      if (d_xl > 0. && d_xl < dist && on[0] != -1) { dist = d_xl; key = -1; }
      if (d_xh > 0. && d_xh < dist && on[0] != 1) { dist = d_xh; key = 1; }
      if (d_yl > 0. && d_yl < dist && on[1] != -1) { dist = d_yl; key = -2; }
      if (d_yh > 0. && d_yh < dist && on[1] != 1) { dist = d_yh; key = 2; }
      if (d_zl > 0. && d_zl < dist && on[2] != -1) { dist = d_zl; key = -3; }
      if (d_zh > 0. && d_zh < dist && on[2] != 1) { dist = d_zh; key = 3; }
      if (dist == ABX_INF) break;
      const double d_tile = dist < remaining ? dist : remaining;
      if (i >= 0 && i < T.Nx && j >= 0 && j < T.Ny && k >= 0 && k < T.Nz) {
        s.tally_add(data + ((int64_t)i * T.Ny + j) * T.Nz + k, d_tile * base);
      } else {
        break;
      }
      remaining -= d_tile;
      if (remaining <= 0.) break;
      r = vadd(r, vscale(u, d_tile));
      on[0] = on[1] = on[2] = 0;
      switch (key) {
        case -1: i--; on[0] = 1; break;
        case 1: i++; on[0] = -1; break;
        case -2: j--; on[1] = 1; break;
        case 2: j++; on[1] = -1; break;
        case -3: k--; on[2] = 1; break;
        case 3: k++; on[2] = -1; break;
        default: break;
      }
    }
  }
}
#endif

struct Scores6 {
  double k_col, k_abs, k_trk, k_tot, leak, mig;
};

// Transporter::collision + branching_collision.  `kcol` is Tallies::kcol().
template <class Sink>
ABX_HD void collide(const DevProblem& P, Sink& s, Part& p, int m, double kcol,
                    bool converged, Scores6& sc) {
  constexpr bool SPLIT = Sink::split;
  const int G = P.G;
  const int mg = m * G + p.g;
  if (converged && P.ntally) score_collision_tallies(P, s, p, m);
  const double total = P.Et[mg], fission = P.Ef[mg], absorption = P.Eabs[mg];
  const double nu_total = P.nu_tot[mg], nu_delayed = P.nu_d[mg];
  // ratios to the total cross section: nu-fission, absorption, nu * fission
  double r_vf, r_abs, r_nuf;
  if constexpr (SPLIT) {
    const float it = fp_rcp((float)total);
    r_vf = (double)((float)P.vEf[mg] * it);
    r_abs = (double)((float)absorption * it);
    r_nuf = (double)((float)(nu_total * fission) * it);
  }
  {
    const V3 dr = vsub(p.r, p.rb);
    if constexpr (SPLIT) {
      sc.k_col += p.w * r_vf;
      sc.mig += p.w * r_abs * vdot(dr, dr);
    } else {
      const double k_col_scr = p.w * P.vEf[mg] / P.Et[mg];
      const double mig_dist = vnorm(dr);
      const double mig_area_scr = p.w * P.Eabs[mg] / P.Et[mg] * mig_dist * mig_dist;
      sc.k_col += k_col_scr;
      sc.mig += mig_area_scr;
    }
  }
  (void)p_rand(p);  // MaterialHelper::sample_nuclide (single component)
  double k_abs_scr;
  if constexpr (SPLIT)
    k_abs_scr = p.w * r_nuf;
  else
    k_abs_scr = p.w * nu_total * fission / total;
  sc.k_abs += k_abs_scr;
  // make_fission_neutrons
  {
    int n_new;
    if (P.mode == 0)
      n_new = (int)floor(dabs(k_abs_scr) / kcol + p_rand(p));
    else
      n_new = (int)floor(dabs(k_abs_scr) + p_rand(p));
    const double P_delayed = nu_delayed / nu_total;
    const double* chi = P.chi_cdf + (size_t)mg * G;
    for (int i = 0; i < n_new; i++) {
      const int ei = p_discrete(p, chi, G);
      const double E_out = P.emid[ei];
      const double mu = 2. * p_rand(p) - 1.;
      const double phi = 2. * ABX_PI * p_rand(p);
      const V3 uout = rotate<SPLIT>(p.u, mu, phi);
      if (p_rand(p) < P_delayed) {
        (void)p_discrete(p, P.dgrp_cdf + (size_t)m * P.n_dgrp_max, P.dgrp_n[m]);
      }
      const double wgt = p.w > 0. ? 1. : -1.;
      if (P.mode == 1) {
        SecRec q{{p.r.x, p.r.y, p.r.z}, {uout.x, uout.y, uout.z}, E_out, wgt, 0.};
        if (!s.sec_push(q)) {
          s.error(-5, p.hid);
          p.alive = 0;
          return;
        }
      } else {
        s.bank_fission(p, uout, E_out, wgt);
        p.daughters++;
      }
    }
  }
  if constexpr (SPLIT) {
    const double keep = 1. - r_abs;
    p.w = p.w * keep;
    p.w2 = p.w2 * keep;
  } else {
    p.w = p.w * (1. - (absorption + 0.) / total);
    p.w2 = p.w2 * (1. - (absorption + 0.) / total);
  }
  // russian_roulette
  if (dabs(p.w) < P.wgt_cutoff) {
    const double P_kill = 1.0 - (dabs(p.w) / P.wgt_survival);
    if (p_rand(p) < P_kill)
      p.w = 0.;
    else
      p.w = copysign(P.wgt_survival, p.w);
  }
  if (dabs(p.w2) < P.wgt_cutoff) {
    const double P_kill = 1.0 - (dabs(p.w2) / P.wgt_survival);
    if (p_rand(p) < P_kill)
      p.w2 = 0.;
    else
      p.w2 = copysign(P.wgt_survival, p.w2);
  }
  if (p.w == 0. && p.w2 == 0.) p.alive = 0;
  if (p.alive) {
    // do_scatter -> MGNuclide::sample_scatter
    const int ei = p_discrete(p, P.scat_cdf + (size_t)mg * G, G);
    const double E_out = P.emid[ei];
    const double mu = sample_mu(P, p, m, p.g, ei);
    const double phi = 2. * ABX_PI * p_rand(p);
    if constexpr (Sink::ext) {  // Particle::set_direction / set_energy
      p.pu = p.u;
      p.pE = p.E;
    }
    p.u = rotate<SPLIT>(p.u, mu, phi);
    p.E = E_out;
    p.g = energy_index(P, E_out);
    p.w = p.w * 1.;
    p.w2 = p.w2 * 1.;
    if (p.E < P.min_energy) p.alive = 0;
  }
}

template <class Sink>
ABX_HD void leak(Part& p, const Bound& b, Scores6& sc) {
  p.alive = 0;
  sc.leak += p.w;
  const V3 r_leak = vadd(p.r, vscale(p.u, b.d));
  const V3 dr = vsub(r_leak, p.rb);
  sc.mig += vdot(vscale(dr, p.w), dr);
}

// Tracker::do_reflection; the particle's position before the reflection
// only feeds Particle::previous_r, which this path does not keep.
// EXT: also the previous position the reference computes for exact
// cancellation (tracker.hpp:328-356): the point as far behind the surface,
// along the new direction, as the last sampled position was before it.
template <bool EXT, class TK>
ABX_HD void reflect(const DevProblem& P, TK& t, Part& p, const Bound& b) {
  if (b.sidx < 0) {
    t.err = -3;
    return;
  }
  const DSurf& S = P.surf[b.sidx];
  const V3 r_on = vadd(p.r, vscale(p.u, b.d));
  const V3 n = surf_norm<TK::kExt>(S, r_on);
  const V3 nd = vsub(p.u, vscale(n, 2. * vdot(p.u, n)));
  const V3 u_new = vdir(nd.x, nd.y, nd.z);
  if constexpr (EXT) {
    const V3 r_pre = p.refl ? p.pr : p.r;
    const double d = b.d + vnorm(vsub(p.r, r_pre));
    p.pr = vsub(r_on, vscale(u_new, d));
    p.pu = p.u;
    p.refl = 1;
  }
  p.r = r_on;
  p.u = u_new;
  t.token = 0;
  t.set_lv(0, p.r);
  trk_restart(P, t, p.u);
}

template <class Sink, class TK>
ABX_HD bool try_resurrect(const DevProblem& P, Sink& s, Part& p, TK& t, int& m) {
  SecRec q;
  if (!s.sec_pop(q)) return false;
  p.alive = 1;
  p.r = V3{q.r[0], q.r[1], q.r[2]};
  p.u = V3{q.u[0], q.u[1], q.u[2]};
  p.E = q.E;
  p.g = energy_index(P, p.E);
  p.w = q.w;
  p.w2 = q.w2;
  p.sec_id++;
  t.token = 0;
  t.set_lv(0, p.r);
  trk_restart(P, t, p.u);
  if (t.cell < 0 || t.err) {
    s.error(-3, p.hid);
    p.alive = 0;
    return false;
  }
  m = P.cell[t.cell].fill;
  s.count_secondary();
  s.trace(p, t, 6);
  return true;
}

// Carter-tracking weight split (src/carter_tracker.cpp:219-223)
template <class Sink>
ABX_HD void split_weight(Sink& s, Part& p) {
  const int n_new = (int)ceil(dabs(p.w));
  if (n_new > 1) {
    p.w = p.w / (double)n_new;
    p.w2 = p.w2 / (double)n_new;
    for (int np = 0; np < n_new - 1; np++) {
      SecRec q{{p.r.x, p.r.y, p.r.z}, {p.u.x, p.u.y, p.u.z}, p.E, p.w, p.w2};
      if (!s.sec_push(q)) {
        s.error(-5, p.hid);
        p.alive = 0;
        return;
      }
    }
  }
}

template <class Sink>
ABX_HD void split(Sink& s, Part& p, const DevProblem& P) {
  if (p.alive && dabs(p.w) >= P.wgt_split) split_weight(s, p);
}

// Start of a history: locate; returns false if the particle is dead
// (lost at birth, DeltaTracker::transport src/delta_tracker.cpp:88-98).
template <class Sink, class TK>
ABX_HD bool start_history(const DevProblem& P, Sink& s, Part& p, TK& t, int& m) {
  trk_init(P, t, p.r, p.u, 0);
  if (t.err) {
    s.error(-3, p.hid);
    p.alive = 0;
    return false;
  }
  if (t.cell < 0) {
    p.alive = 0;
    s.count_lost();
    return false;
  }
  m = P.cell[t.cell].fill;
  return true;
}

// One flight of the transport loop up to the collision decision.  Returns
// true when a real collision is pending; otherwise `ev` holds the event.
// Sets p.alive = 0 and returns false after raising an error.
template <int TRK, class Sink, class TK>
ABX_HD bool step_flight(const DevProblem& P, Sink& s, Part& p, TK& t, int& m, bool converged,
                        Scores6& sc, int& ev) {
  if (TRK == 0) {
    // ---------------- surface tracking (src/surface_tracker.cpp:70-151)
    const double d_coll = flight_distance<Sink::split>(p, P.Et[m * P.G + p.g]);
    s.count_flight();
    const Bound b = trk_nearest(P, t, p.u);
    const double dmin = d_coll < b.d ? d_coll : b.d;
    if (converged && P.ntally) score_flight_tallies(P, s, p, dmin, m);
    sc.k_trk += p.w * dmin * P.vEf[m * P.G + p.g];
    ev = 2;
    if (b.d < d_coll || dabs(b.d - d_coll) < ABX_BOUNDRY_TOL) {
      s.count_crossing();
      if (b.btype == 0) {
        leak<Sink>(p, b, sc);
        ev = 4;
      } else if (b.btype == 1) {
        reflect<Sink::ext>(P, t, p, b);
        if (t.cell < 0 || t.err) {
          s.error(-3, p.hid);
          p.alive = 0;
          return false;
        }
        ev = 3;
      } else {
        trk_move(t, p.u, b.d);
        t.token = -b.tok;
        trk_get_current(P, t, p.u);
        part_move<Sink::ext>(p, b.d);
        if (t.cell < 0 || t.err) {
          s.error(-3, p.hid);
          p.alive = 0;
          return false;
        }
        m = P.cell[t.cell].fill;
      }
      return false;
    }
    part_move<Sink::ext>(p, d_coll);
    trk_move(t, p.u, d_coll);
    return true;
  }
  // ---------------- delta / carter tracking (src/delta_tracker.cpp:105-195,
  // src/carter_tracker.cpp:124-216)
  const bool carter = TRK == 2;
  const double Emaj = (carter ? P.sampling[p.g] : P.majorant[p.g]) + 0.;
  if constexpr (Sink::ext) p.Esmp = Emaj;  // Particle::set_Esmp
  const double d_coll = flight_distance<Sink::split>(p, Emaj);
  s.count_flight();
  trk_move(t, p.u, d_coll);
  trk_get_current(P, t, p.u);
  if (t.err) {
    s.error(-3, p.hid);
    p.alive = 0;
    return false;
  }
  const bool lost = t.cell < 0;
  Bound b{ABX_INF, -1, 2, 0};
  if (lost) {
    trk_reset_position(P, t, p.r, p.u);
    if (t.err) {
      s.error(-3, p.hid);
      p.alive = 0;
      return false;
    }
    b = trk_bc_cold(P, t, p.u);
  }
  if (converged && P.ntally) score_flight_tallies(P, s, p, d_coll < b.d ? d_coll : b.d, m);
  if (lost) {
    s.count_crossing();
    if (b.btype == 0) {
      leak<Sink>(p, b, sc);
      ev = 4;
    } else if (b.btype == 1) {
      reflect<Sink::ext>(P, t, p, b);
      if (t.cell < 0 || t.err) {
        s.error(-3, p.hid);
        p.alive = 0;
        return false;
      }
      ev = 3;
    } else {
      s.error(-3, p.hid);
      p.alive = 0;
    }
    return false;
  }
  ev = 0;
  part_move<Sink::ext>(p, d_coll);
  m = P.cell[t.cell].fill;
  const double Et = P.Et[m * P.G + p.g];
  bool had_collision = false;
  if (!carter) {
    if (Et - Emaj > 1.E-10) {
      s.error(-4, p.hid);
      p.alive = 0;
      return false;
    }
    if constexpr (Sink::split) {
      if ((float)p_rand(p) < fp_div((float)Et, (float)Emaj)) had_collision = true;
    } else {
      if (p_rand(p) < (Et / Emaj)) had_collision = true;
    }
  } else if (Emaj >= Et) {
    if (p_rand(p) < (Et / Emaj)) had_collision = true;
  } else {
    const double D = Et / (2. * Et - Emaj);
    const double F = Et / (D * Emaj);
    if ((D - p_rand(p)) > 0.) {
      p.w = p.w * F;
      had_collision = true;
    } else {
      p.w = -p.w * F;
    }
  }
  if (!had_collision) s.count_virtual();
  return had_collision;
}

// Real collision; sets the event code.
template <class Sink>
ABX_HD void step_collision(const DevProblem& P, Sink& s, Part& p, int m, double kcol,
                           bool converged, Scores6& sc, int& ev) {
  s.count_collision();
  collide(P, s, p, m, kcol, converged, sc);
  ev = p.alive ? 1 : 5;
}

// End of a loop iteration: carter split, trace, resurrection.  Returns
// whether the slot still carries a live particle.
template <int TRK, class Sink, class TK>
ABX_HD bool step_finish(const DevProblem& P, Sink& s, Part& p, TK& t, int& m, int ev) {
  if constexpr (Sink::ext && TRK != 0) {
    // set_previous_collision_real / _virtual (src/delta_tracker.cpp:187-195,
    // src/carter_tracker.cpp:209-217): real after a collision, virtual
    // after any other event the particle survives (reflections included)
    if (ev == 1 || ev == 5)
      p.pvirt = 0;
    else if (p.alive)
      p.pvirt = 1;
  }
  if (TRK == 2) split(s, p, P);
  s.trace(p, t, ev);
  if (!p.alive) try_resurrect(P, s, p, t, m);
  return p.alive != 0;
}

// One whole history (plus its secondaries).
template <int TRK, class Sink, class TK>
ABX_HD void run_loop(const DevProblem& P, Sink& s, Part& p, TK& t, int& m, double kcol,
                     bool converged, Scores6& sc, bool collision_pending) {
  for (int guard = 0; p.alive; guard++) {
    if (guard > 100000000) {
      s.error(-3, p.hid);
      return;
    }
    int ev = 0;
    bool hc = collision_pending;
    collision_pending = false;
    if (!hc) hc = step_flight<TRK>(P, s, p, t, m, converged, sc, ev);
    if (p.alive && hc) step_collision(P, s, p, m, kcol, converged, sc, ev);
    if (!step_finish<TRK>(P, s, p, t, m, ev)) return;
  }
}

template <int TRK, class Sink, class TK>
ABX_HD void run_history(const DevProblem& P, Sink& s, Part& p, TK& t, double kcol,
                        bool converged, Scores6& sc) {
  int m = -1;
  if (!start_history(P, s, p, t, m)) return;
  run_loop<TRK>(P, s, p, t, m, kcol, converged, sc, false);
}

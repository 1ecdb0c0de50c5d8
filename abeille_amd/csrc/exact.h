// exact.h - the per-site and per-region arithmetic of the exact multigroup
// cancelators (ExactMGCancelator, src/exact_mg_cancelator.cpp, and
// BasicExactMGCancelator, src/basic_exact_mg_cancelator.cpp), shared by the
// device kernels (exact_kernels.h) and host builds of the device code.
//
// A cancellation region is (i, j, k, e) of the cancellation mesh and a
// material; its sites are the fission sites of a generation that fall in it.
// Regions are numbered so that sorting their numbers orders them by
// (i, j, k, e) lexicographically, then by material: the reference keeps them
// in unordered maps (by Key and by Material*), whose iteration order is the
// library's; this is the fixed order every draw from the global pcg32 follows.
#pragma once

#include <cmath>
#include <cstdint>

#include "../../include/abx.h"
#include "physics.h"

// Sobol generator matrices of the first four dimensions of the Joe-Kuo
// sequence (new-joe-kuo-6.21201) with 52 bits, as sobol::sample uses them
// (vendor/sobol/include/sobol/sobol.hpp): dimension 0 is the van der Corput
// sequence; dimensions 1-3 come from the primitive polynomials of degree
// s = 1, 2, 3 (a = 0, 1, 1) and the initial direction numbers m = {1},
// {1, 3}, {1, 3, 1}, extended by Bratley and Fox's recurrence.
constexpr int kSobolBits = 52;
constexpr int kSobolDims = 4;

static inline void sobol_build(uint64_t* M) {
  for (int k = 0; k < kSobolBits; k++) M[k] = 1ull << (kSobolBits - 1 - k);
  const int deg[3] = {1, 2, 3}, coef[3] = {0, 1, 1};
  const uint64_t m0[3][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
  for (int d = 0; d < 3; d++) {
    uint64_t* V = M + (d + 1) * kSobolBits;
    const int s = deg[d];
    for (int k = 0; k < s; k++) V[k] = m0[d][k] << (kSobolBits - 1 - k);
    for (int k = s; k < kSobolBits; k++) {
      uint64_t v = V[k - s] ^ (V[k - s] >> s);
      for (int q = 1; q < s; q++)
        if ((coef[d] >> (s - 1 - q)) & 1) v ^= V[k - q];
      V[k] = v;
    }
  }
}

// sobol::sample(index, dim): the index's bits select matrix columns, XORed,
// scaled by 2^-52 (exact in FP64)
ABX_HD double sobol_at(const uint64_t* M, uint64_t index, int dim) {
  const uint64_t* V = M + dim * kSobolBits;
  uint64_t r = 0;
  for (int i = 0; index != 0; index >>= 1, i++)
    if (index & 1u) r ^= V[i];
  return (double)r * (1.0 / 4503599627370496.0);
}

// Device view of the cancelator configuration.
struct DevExact {
  int32_t kind;         // abx_exact_kind
  int32_t beta;         // abx_beta_mode (basic)
  int32_t n_samples;
  int32_t chi_matrix;
  int32_t use_virtual;
  int32_t sobol;        // basic: Sobol points (else the pcg32)
  int32_t S[4];         // regions in x, y, z; energy regions (1 without group bins)
  int32_t n_gbins;      // group bins given (0 = none)
  int32_t nmat;
  int32_t G;
  int32_t pad_;
  double low[3], hi[3], pitch[3];
  const int32_t* gb_off;  // [n_gbins + 1]
  const int32_t* gb;      // the groups of each bin
  const int32_t* g2bin;   // [G] the bin holding group g, n_gbins if none
  const double* chi;      // [nmat][G][G] normalised fission spectra (chi_matrix)
  const uint64_t* sobolM; // [4][52]
};

// settings::group (src/settings.cpp:95-104): the first group whose bounds
// hold E, inclusive at both ends; 0 when none does
ABX_HD int ex_group(const DevProblem& P, double E) {
  for (int g = 0; g < P.G; g++)
    if (P.ebounds[g] <= E && E <= P.ebounds[g + 1]) return g;
  return 0;
}

// Region number of (i, j, k, e) on the mesh widened by one in every
// direction (i, j, k may equal the shape on the mesh's upper faces, e the
// number of group bins for a group in none), times the materials, plus the
// material.  kExNone marks a site in no region.
constexpr uint32_t kExNone = 0xffffffffu;

ABX_HD uint64_t ex_region(const DevExact& E, uint32_t i, uint32_t j, uint32_t k, uint32_t e,
                          int mat) {
  const uint64_t ijke =
      (uint64_t)e + (uint64_t)(E.S[3] + 1) *
                        ((uint64_t)k + (uint64_t)(E.S[2] + 1) *
                                           ((uint64_t)j + (uint64_t)(E.S[1] + 1) * (uint64_t)i));
  return ijke * (uint64_t)E.nmat + (uint64_t)mat;
}

struct ExKeyD {
  uint32_t i, j, k, e;
  int32_t mat;
};

ABX_HD ExKeyD ex_decode(const DevExact& E, uint64_t reg) {
  ExKeyD K;
  K.mat = (int32_t)(reg % (uint64_t)E.nmat);
  uint64_t q = reg / (uint64_t)E.nmat;
  K.e = (uint32_t)(q % (uint64_t)(E.S[3] + 1));
  q /= (uint64_t)(E.S[3] + 1);
  K.k = (uint32_t)(q % (uint64_t)(E.S[2] + 1));
  q /= (uint64_t)(E.S[2] + 1);
  K.j = (uint32_t)(q % (uint64_t)(E.S[1] + 1));
  K.i = (uint32_t)(q / (uint64_t)(E.S[1] + 1));
  return K;
}

// The region indices of a site (ExactMGCancelator::get_key, :117-152 and
// add_particle, :172-209; BasicExactMGCancelator::add_particle, :68-113);
// false when the site joins no region.
ABX_HD bool ex_site_key(const DevExact& E, const DevProblem& P, const abx_banked& b,
                        int32_t prev_virtual, ExKeyD& K) {
  const double x = b.r[0], y = b.r[1], z = b.r[2];
  if (E.kind == ABX_EXACT_MG) {
    if (x < E.low[0] || x > E.hi[0] || y < E.low[1] || y > E.hi[1] || z < E.low[2] ||
        z > E.hi[2])
      return false;
    K.i = (uint32_t)floor((x - E.low[0]) / E.pitch[0]);
    K.j = (uint32_t)floor((y - E.low[1]) / E.pitch[1]);
    K.k = (uint32_t)floor((z - E.low[2]) / E.pitch[2]);
    const int g = ex_group(P, b.E);
    K.e = (uint32_t)E.g2bin[g];
    if (K.e == (uint32_t)E.n_gbins && E.chi_matrix) return false;
    // a site right after a virtual collision has a region but is not added
    if (!E.use_virtual && prev_virtual) return false;
    return true;
  }
  const int i = (int)floor((x - E.low[0]) / E.pitch[0]);
  const int j = (int)floor((y - E.low[1]) / E.pitch[1]);
  const int k = (int)floor((z - E.low[2]) / E.pitch[2]);
  if (i < 0 || j < 0 || k < 0) return false;
  if (i >= E.S[0] || j >= E.S[1] || k >= E.S[2]) return false;
  K.i = (uint32_t)i;
  K.j = (uint32_t)j;
  K.k = (uint32_t)k;
  K.e = 0;
  return true;
}

// get_material (:154-170): the material of the cell holding r, located with
// u = (1, 0, 0); -1 in no cell
template <class TK>
ABX_HD int ex_material(const DevProblem& P, TK& t, const V3& r) {
  trk_init(P, t, r, V3{1., 0., 0.}, 0);
  if (t.err || t.cell < 0) return -1;
  return P.cell[t.cell].fill;
}

// MGAngleDistribution::pdf (include/materials/mg_angle_distribution.hpp:62-75)
// of the scattering-cosine distribution g1 -> g3 of material m; a material
// without tables has the isotropic {-1, 1} / {0.5, 0.5} one, pdf 0.5.
ABX_HD double ex_angle_pdf(const DevProblem& P, int m, int g1, int g3, double mu) {
  const int32_t base = P.ang_base[m];
  if (base < 0) return 0.5;
  const int32_t b0 = P.ang_off[base + g1 * P.G + g3], b1 = P.ang_off[base + g1 * P.G + g3 + 1];
  const double* am = P.ang_mu;
  const double* ap = P.ang_pdf;
  if (mu < am[b0]) return ap[b0];
  if (mu > am[b1 - 1]) return ap[b1 - 1];
  int lo = b0, hi = b1 - 1;  // first entry not below mu (lower_bound)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (am[mid] < mu)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (mu == am[lo]) return ap[lo];
  const int l = lo - 1;
  const double sl = (ap[l + 1] - ap[l]) / (am[l + 1] - am[l]);
  return sl * (mu - am[l]) + ap[l];
}

// ExactMGCancelator::get_f (:220-241): the transition density from the
// parent's previous collision (r1, u1, groups g1 -> g3) to r4 in group g4
ABX_HD double ex_exact_f(const DevExact& E, const DevProblem& P, int m, const V3& r1,
                         const V3& u1, int g1, int g3, const V3& r4, int g4, double Esmp) {
  const double dx = r4.x - r1.x, dy = r4.y - r1.y, dz = r4.z - r1.z;
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  const double nn = sqrt(dx * dx + dy * dy + dz * dz);
  const double ux = dx / nn, uy = dy / nn, uz = dz / nn;
  const double mu = ux * u1.x + uy * u1.y + uz * u1.z;
  const double pdf_mu = ex_angle_pdf(P, m, g1, g3, mu);
  const double pdf_chi = E.chi_matrix ? E.chi[((size_t)m * E.G + g3) * E.G + g4] : 1.;
  return (pdf_mu * pdf_chi / (d * d)) * dm_exp(-Esmp * d);
}

// BasicExactMGCancelator::get_f (:188-216)
ABX_HD double ex_basic_f(const V3& r, const V3& rp, double Esmp) {
  const double dx = r.x - rp.x, dy = r.y - rp.y, dz = r.z - rp.z;
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  return (1. / (d * d)) * dm_exp(-Esmp * d);
}

ABX_HD double ex_min2(double a, double b) { return b < a ? b : a; }  // std::min

// BasicExactMGCancelator::get_min_f (:218-241): the smallest f over the
// region's corners
ABX_HD double ex_basic_min_f(const DevExact& E, const ExKeyD& K, const V3& rp, double Esmp) {
  const double Xl = E.low[0] + (double)K.i * E.pitch[0], Xh = Xl + E.pitch[0];
  const double Yl = E.low[1] + (double)K.j * E.pitch[1], Yh = Yl + E.pitch[1];
  const double Zl = E.low[2] + (double)K.k * E.pitch[2], Zh = Zl + E.pitch[2];
  const double f1 = ex_basic_f({Xh, Yh, Zh}, rp, Esmp), f2 = ex_basic_f({Xh, Yh, Zl}, rp, Esmp);
  const double f3 = ex_basic_f({Xh, Yl, Zh}, rp, Esmp), f4 = ex_basic_f({Xh, Yl, Zl}, rp, Esmp);
  const double f5 = ex_basic_f({Xl, Yh, Zh}, rp, Esmp), f6 = ex_basic_f({Xl, Yh, Zl}, rp, Esmp);
  const double f7 = ex_basic_f({Xl, Yl, Zh}, rp, Esmp), f8 = ex_basic_f({Xl, Yl, Zl}, rp, Esmp);
  return ex_min2(ex_min2(ex_min2(f1, f2), ex_min2(f3, f4)),
                 ex_min2(ex_min2(f5, f6), ex_min2(f7, f8)));
}

// A point of region K from three uniform numbers (sample_point /
// get_averages' candidate positions)
ABX_HD V3 ex_point(const DevExact& E, const ExKeyD& K, double xi0, double xi1, double xi2) {
  const double Xl = E.low[0] + (double)K.i * E.pitch[0];
  const double Yl = E.low[1] + (double)K.j * E.pitch[1];
  const double Zl = E.low[2] + (double)K.k * E.pitch[2];
  return V3{Xl + xi0 * E.pitch[0], Yl + xi1 * E.pitch[1], Zl + xi2 * E.pitch[2]};
}

// C(f, f_inv) of compute_averages (:343-345)
ABX_HD double ex_c(double f, double f_inv) { return 1. / (2. * f * f_inv - 1.); }

// pcg32 steps one uniform particle of get_new_particles takes after its
// accepted position (three numbers): the exact cancelator draws the energy
// and the direction's two angles; the basic one calls sample_fission, which
// draws the energy, two angles and the delayed-neutron test.  The energy
// takes no number when it is a discrete draw over a single group
// (std::discrete_distribution of one weight draws nothing).
ABX_HD uint64_t ex_tail_steps(const DevExact& E) {
  const uint64_t energy = ((E.kind == ABX_EXACT_MG && E.chi_matrix) || E.G >= 2) ? 1u : 0u;
  return 2u * (energy + (E.kind == ABX_EXACT_MG ? 2u : 3u));
}

// Energy and direction of a uniform particle (get_new_particles, exact
// :552-577; basic :573-600 through MGNuclide::sample_fission,
// src/mg_nuclide.cpp:504-544), from the pcg32 state after its position.
ABX_HD void ex_new_tail(const DevExact& E, const DevProblem& P, const ExKeyD& K, uint64_t& s,
                        abx_banked& u) {
  if (E.kind == ABX_EXACT_MG) {
    int e_index;
    if (E.chi_matrix) {
      const double xi_E = rng_rand(s);
      const int nb = E.gb_off[K.e + 1] - E.gb_off[K.e];
      const int gi = (int)floor(xi_E * (double)nb);
      e_index = E.gb[E.gb_off[K.e] + gi];
    } else {
      e_index = rng_discrete(s, P.chi_cdf + ((size_t)K.mat * P.G) * P.G, P.G);
    }
    u.E = 0.5 * (P.ebounds[e_index] + P.ebounds[e_index + 1]);
    // Direction u_smp(2 rand - 1, 2 pi rand): GCC evaluates the arguments
    // right to left, so phi is drawn first
    const double phi = 2. * ABX_PI * rng_rand(s);
    const double mu = 2. * rng_rand(s) - 1.;
    const V3 d = dir_mu_phi(mu, phi);
    u.u[0] = d.x;
    u.u[1] = d.y;
    u.u[2] = d.z;
  } else {
    const int ei = rng_discrete(s, P.chi_cdf + ((size_t)K.mat * P.G) * P.G, P.G);
    u.E = 0.5 * (P.ebounds[ei] + P.ebounds[ei + 1]);
    const double mu = 2. * rng_rand(s) - 1.;
    const double phi = 2. * ABX_PI * rng_rand(s);
    const V3 d = rotate_dir(V3{0., 0., 1.}, mu, phi);
    u.u[0] = d.x;
    u.u[1] = d.y;
    u.u[2] = d.z;
    (void)rng_rand(s);  // the delayed-neutron test, rand() < Pdelayed = 0
  }
}

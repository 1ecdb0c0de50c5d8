// exact_power_iteration.cpp - PowerIterator::run with regional cancellation
// (src/power_iterator.cpp:305-429, perform_regional_cancellation :751-777)
// written against the C++ host mirror (include/abeille_gpu.hpp):
// GpuTransporter as the Transporter and GpuExactCancelator as the
// Cancelator, drawing from the driver's own settings::rng.  The problem is
// the PUa-1-0-SL slab under carter tracking at 0.85 x the majorant (signed
// weights), with an exact (argv[1] = "exact") or basic-exact average-g
// (argv[1] = "basic") cancelator on 8 x 1 x 1 regions.  Prints k per
// generation of this loop ("kgen"), then of abx_run on the same problem
// ("krun"), which runs the same loop on the device.  Built and run by
// tests/test_cpp_host_api.py.
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/abeille_gpu.hpp"

int main(int argc, char** argv) {
  const bool basic = argc > 1 && std::strcmp(argv[1], "basic") == 0;
  const int N = argc > 2 ? atoi(argv[2]) : 4000;
  const int ngen = argc > 3 ? atoi(argv[3]) : 6, nignored = argc > 4 ? atoi(argv[4]) : 2;
  abx_surface s[6] = {};
  const double xs[6] = {-1.853722, 1.853722, -5, 5, -5, 5};
  for (int i = 0; i < 6; i++) {
    s[i].type = i / 2;
    s[i].boundary = i < 2 ? ABX_VACUUM : ABX_REFLECTIVE;
    s[i].id = (uint32_t)(i + 1);
    s[i].p[0] = xs[i];
  }
  const int32_t rpn[6] = {1, -2, 3, -4, 5, -6};
  abx_cell cell{1, ABX_FILL_MATERIAL, 0, 0, 6};
  const int32_t ucells[1] = {0};
  abx_universe uni{};
  uni.id = 1;
  uni.type = ABX_CELL_UNIVERSE;
  uni.n_cells = 1;
  uni.outer = -1;
  uni.elems_offset = -1;
  const double tot[1] = {0.32640}, ab[1] = {0.101184}, sc[1] = {0.225216}, fi[1] = {0.081600},
               nu[1] = {3.24}, chi[1] = {1.0};
  abx_material mat{};
  mat.id = 1;
  mat.total = tot;
  mat.absorption = ab;
  mat.scatter = sc;
  mat.fission = fi;
  mat.nu_prompt = nu;
  mat.chi = chi;
  abx_source src{};
  src.spatial = ABX_SRC_POINT;
  src.direction = ABX_DIR_ISOTROPIC;
  src.energy = 0.5;
  src.weight = 1.0;
  const double eb[2] = {0., 1.}, ratio[1] = {0.85};
  abx_exact_cancel ec{};
  ec.kind = basic ? ABX_BASIC_EXACT_MG : ABX_EXACT_MG;
  ec.shape[0] = 8;
  ec.shape[1] = 1;
  ec.shape[2] = 1;
  for (int k = 0; k < 3; k++) {
    ec.low[k] = xs[2 * k];
    ec.high[k] = xs[2 * k + 1];
  }
  ec.n_samples = 10;
  ec.use_virtual_collisions = 1;
  ec.beta = ABX_BETA_AVERAGE_G;
  ec.sobol = 1;
  abx_problem_desc d{};
  d.n_surfaces = 6;
  d.surfaces = s;
  d.n_cells = 1;
  d.cells = &cell;
  d.n_rpn = 6;
  d.rpn = rpn;
  d.n_universes = 1;
  d.universes = &uni;
  d.n_universe_cells = 1;
  d.universe_cells = ucells;
  d.n_materials = 1;
  d.materials = &mat;
  d.n_sources = 1;
  d.sources = &src;
  d.settings.mode = ABX_K_EIGENVALUE;
  d.settings.tracking = ABX_CARTER_TRACKING;
  d.settings.seed = 19073486328125ULL;
  d.settings.stride = 152917ULL;
  d.settings.wgt_cutoff = 0.25;
  d.settings.wgt_survival = 1.0;
  d.settings.wgt_split = 2.0;
  d.settings.nparticles = N;
  d.settings.ngroups = 1;
  d.settings.energy_bounds = eb;
  d.settings.sample_xs_ratio = ratio;
  d.exact_cancel = &ec;
  try {
    auto prob = std::make_shared<abx::Problem>(d);
    auto tallies = std::make_shared<abx::Tallies>(prob, (double)N);
    std::vector<abx_particle> src_bank(N);
    abx::check(abx_sample_sources(prob->get(), 0, N, src_bank.data()));
    std::vector<abx::Particle> bank;
    for (auto& a : src_bank) {
      abx::Particle p({a.r[0], a.r[1], a.r[2]}, {a.u[0], a.u[1], a.u[2]}, a.E, a.wgt, a.history_id);
      p.rng_state = a.rng_state;
      p.set_family_id(a.history_id);
      bank.push_back(p);
    }
    uint64_t histories_counter = (uint64_t)N;
    abx::pcg32 rng(d.settings.seed);  // settings::rng
    std::shared_ptr<abx::Transporter> transporter =
        std::make_shared<abx::GpuTransporter>(tallies);
    std::shared_ptr<abx::Cancelator> cancelator = std::make_shared<abx::GpuExactCancelator>(prob);
    if (nignored == 0) abx::settings::converged = true;
    for (int g = 1; g <= ngen; g++) {
      std::vector<abx::BankedParticle> next_gen = transporter->transport(bank);
      if (next_gen.size() == 0) throw abx::Error(ABX_E_INVALID, "No fission neutrons were produced.");
      tallies->calc_gen_values();
      // perform_regional_cancellation (src/power_iterator.cpp:751-777)
      for (auto& p : next_gen) cancelator->add_particle(p);
      cancelator->perform_cancellation(rng);
      auto tmp = cancelator->get_new_particles(rng);
      next_gen.insert(next_gen.end(), tmp.begin(), tmp.end());
      cancelator->clear();
      // normalize_weights (src/power_iterator.cpp:538-586)
      double W_pos = 0., W_neg = 0.;
      for (auto& b : next_gen) {
        if (b.wgt > 0.)
          W_pos += b.wgt;
        else
          W_neg -= b.wgt;
      }
      const double w_per_part = (double)N / (W_pos - W_neg);
      for (auto& b : next_gen) b.wgt *= w_per_part;
      if (abx::settings::converged) tallies->record_generation();
      tallies->clear_generation();
      bank.clear();
      for (auto& b : next_gen) {
        bank.push_back(abx::Particle(b.r, b.u, b.E, b.wgt, histories_counter++));
        bank.back().initialize_rng(d.settings.seed, d.settings.stride);
        bank.back().set_family_id(b.family_id);
      }
      std::printf("kgen %d %.17g %zu\n", g, tallies->kcol(), next_gen.size());
      if (g == nignored) abx::settings::converged = true;
    }
    // the same loop on the device
    std::vector<double> kh(ngen);
    abx_run_stats st{};
    abx::check(abx_run(prob->get(), ngen, nignored, kh.data(), &st));
    for (int g = 0; g < ngen; g++) std::printf("krun %d %.17g\n", g + 1, kh[g]);
  } catch (const abx::Error& e) {
    std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
